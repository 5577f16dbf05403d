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
//              oddTextureName/evenTextureName), image (name filepath: PNG)
//   materials  lambertian (name color | texture), metal (name color fuzziness), glass
//              (name refraction) and dielectric (same keys)
//   primitives spheres (name position radius material uvmapping), quads (name position u v
//              material), boxes (name position size material)
//   meshes     obj (name filepath material use_smoothing active)
// Semantics kept from the reference loader: colours written as integers are 0-255
// (getColor, scene_loader.cpp:1622-1641), metal reads `fuzziness` only (scene_loader.cpp:1128),
// box `position` is the centre (core/src/primitives/box.cuh:41-46), quad `position` is a corner,
// directional-light `position` is the centre. Fixed on purpose: counts are not truncated to 8 bits
// and `glass` materials are loaded (the reference sizes them from the `dielectric` key,
// scene_loader.cpp:191-192 vs 1092-1108, and drops them otherwise). Transforms, groups and the
// features outside the render path (omni/spot lights, other materials, textures and primitives)
// are counted in `unsupported` and skipped.
#include <zlib.h>

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

// ------------------------------------------------------------------ PNG reader (8-bit, zlib inflate)
static bool read_png(const std::string& path, std::vector<uint8_t>& out, uint32_t& W, uint32_t& H, uint32_t& C,
                     std::string& err) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) { err = "cannot open " + path; return false; }
    std::vector<uint8_t> d;
    uint8_t buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) d.insert(d.end(), buf, buf + n);
    std::fclose(f);
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    if (d.size() < 8 || std::memcmp(d.data(), sig, 8) != 0) { err = path + ": not a PNG file (only PNG textures are read)"; return false; }
    auto be32 = [&](size_t o) { return (uint32_t)d[o] << 24 | (uint32_t)d[o + 1] << 16 | (uint32_t)d[o + 2] << 8 | d[o + 3]; };
    size_t o = 8;
    uint32_t depth = 0, ctype = 0, interlace = 0;
    std::vector<uint8_t> idat, pal;
    W = H = 0;
    while (o + 8 <= d.size()) {
        uint32_t len = be32(o);
        std::string tag((const char*)&d[o + 4], 4);
        if (o + 12 + (size_t)len > d.size()) { err = path + ": truncated chunk"; return false; }
        const uint8_t* p = &d[o + 8];
        if (tag == "IHDR") { W = be32(o + 8); H = be32(o + 12); depth = p[8]; ctype = p[9]; interlace = p[12]; }
        else if (tag == "PLTE") pal.assign(p, p + len);
        else if (tag == "IDAT") idat.insert(idat.end(), p, p + len);
        else if (tag == "IEND") break;
        o += 12 + len;
    }
    if (depth != 8 || interlace != 0 || W == 0 || H == 0 || (uint64_t)W * H > (1ull << 28)) {
        err = path + ": only 8-bit non-interlaced PNG is supported";
        return false;
    }
    uint32_t src_c = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
    if (!src_c) { err = path + ": unsupported PNG colour type"; return false; }
    size_t stride = (size_t)W * src_c;
    std::vector<uint8_t> raw((stride + 1) * H);
    uLongf rl = (uLongf)raw.size();
    if (uncompress(raw.data(), &rl, idat.data(), (uLong)idat.size()) != Z_OK || rl != raw.size()) {
        err = path + ": corrupt image data";
        return false;
    }
    std::vector<uint8_t> img(stride * H);
    for (uint32_t y = 0; y < H; ++y) {
        uint8_t ft = raw[y * (stride + 1)];
        const uint8_t* s = &raw[y * (stride + 1) + 1];
        uint8_t* r = &img[y * stride];
        const uint8_t* up = y ? &img[(y - 1) * stride] : nullptr;
        for (size_t x = 0; x < stride; ++x) {
            int a = x >= src_c ? r[x - src_c] : 0, b = up ? up[x] : 0, c = (up && x >= src_c) ? up[x - src_c] : 0;
            int v = s[x];
            switch (ft) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += b; break;
                case 3: v += (a + b) / 2; break;
                case 4: {
                    int pp = a + b - c, pa = std::abs(pp - a), pb = std::abs(pp - b), pc = std::abs(pp - c);
                    v += (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
                    break;
                }
                default: err = path + ": bad PNG filter"; return false;
            }
            r[x] = (uint8_t)v;
        }
    }
    C = (src_c == 2 || src_c == 4) ? 4 : 3;
    out.resize((size_t)W * H * C);
    for (size_t k = 0; k < (size_t)W * H; ++k) {
        uint8_t rgba[4] = {0, 0, 0, 255};
        if (ctype == 3) {
            size_t pi = (size_t)img[k] * 3;
            if (pi + 2 < pal.size()) { rgba[0] = pal[pi]; rgba[1] = pal[pi + 1]; rgba[2] = pal[pi + 2]; }
        } else if (src_c <= 2) {
            rgba[0] = rgba[1] = rgba[2] = img[k * src_c];
            if (src_c == 2) rgba[3] = img[k * 2 + 1];
        } else {
            for (uint32_t q = 0; q < src_c; ++q) rgba[q] = img[k * src_c + q];
        }
        std::memcpy(&out[k * C], rgba, C);
    }
    return true;
}

// ------------------------------------------------------------------ scene mapping
struct SceneFile {
    std::string path, dir;
    crt_scenefile_info info{};
    std::vector<crt_material> mats;
    std::vector<crt_texture> texs;
    std::vector<std::vector<uint8_t>> tex_pixels;
    std::vector<crt_area_light> lights;
    std::vector<crt_sphere> spheres;
    std::vector<crt_quad> quads;
    std::vector<crt_mesh> meshes;
    std::vector<crt_obj_mesh> objs;
    std::vector<std::vector<uint32_t>> box_idx;
    std::vector<std::vector<float>> box_pos;
    std::map<std::string, int32_t> mat_by_name, tex_by_name;
    crt_scene_desc desc{};
    ~SceneFile() {
        for (auto& o : objs) crt_obj_free(&o);
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
    // camera (scene_loader.cpp:335-366)
    crt_camera cam{};
    cam.vup[1] = 1.0f; cam.vfov_deg = 90.0f; cam.focus_dist = 10.0f; cam.time0 = 0.0f; cam.time1 = 1.0f;
    cam.aspect_ratio = I.aspect_ratio;
    if (const cfg::Value* c = root.get("camera")) {
        const cfg::Value* ar = c->get("aspectRatio");
        if (ar && ar->kind == cfg::Value::STRING) {
            double a = 0, b = 0;
            if (std::sscanf(ar->s.c_str(), "%lf:%lf", &a, &b) == 2 && a > 0 && b > 0) cam.aspect_ratio = (float)(a / b);
        } else if (ar) {
            cam.aspect_ratio = num_of(ar, cam.aspect_ratio);
        }
        vec_of(c->get("lookFrom"), cam.lookfrom);
        vec_of(c->get("lookAt"), cam.lookat);
        if (c->get("upAxis")) vec_of(c->get("upAxis"), cam.vup);
        cam.aperture = num_of(c->get("aperture"), 0.0f);
        cam.focus_dist = num_of(c->get("focus"), cam.focus_dist);
        cam.vfov_deg = num_of(c->get("fov"), cam.vfov_deg);
        if (bool_of(c->get("orthographic"), false)) I.unsupported++;
    }
    I.aspect_ratio = cam.aspect_ratio;
    // lights
    if (const cfg::Value* L = root.get("lights")) {
        if (const cfg::Value* dl = list_of(L, "directionals")) {
            for (auto& it : dl->items) {
                if (!bool_of(it->get("active"), true)) continue;
                crt_area_light a{};
                vec_of(it->get("position"), a.center);
                vec_of(it->get("u"), a.u);
                vec_of(it->get("v"), a.v);
                a.intensity = num_of(it->get("intensity"), 0.0f);
                color_of(it->get("color"), a.rgb);
                a.invisible = bool_of(it->get("invisible"), true) ? 1u : 0u;
                sf.lights.push_back(a);
            }
        }
        for (const char* k : {"omnis", "spots"})
            if (const cfg::Value* l = list_of(L, k)) I.unsupported += (uint32_t)l->items.size();
    }
    // textures
    if (const cfg::Value* T = root.get("textures")) {
        auto add = [&](const std::string& name, const crt_texture& t) {
            if (!name.empty()) sf.tex_by_name[name] = (int32_t)sf.texs.size();
            sf.texs.push_back(t);
            sf.tex_pixels.emplace_back();
        };
        if (const cfg::Value* l = list_of(T, "solidColor")) {
            for (auto& it : l->items) {
                crt_texture t{};
                t.type = CRT_TEX_SOLID; t.even = t.odd = -1;
                color_of(it->get("color"), t.color);
                add(str_of(it->get("name")), t);
            }
        }
        if (const cfg::Value* l = list_of(T, "image")) {
            for (auto& it : l->items) {
                std::string fp = resolve(sf, str_of(it->get("filepath")));
                crt_texture t{};
                t.type = CRT_TEX_IMAGE; t.even = t.odd = -1;
                std::vector<uint8_t> px;
                std::string e2;
                if (!read_png(fp, px, t.width, t.height, t.channels, e2)) {
                    // the reference prints and continues with an empty image (core/src/gpu.cu:1744-1748)
                    std::fprintf(stderr, "[WARNING] %s\n", e2.c_str());
                    t.width = t.height = 0; t.channels = 3;
                }
                add(str_of(it->get("name")), t);
                sf.tex_pixels.back() = std::move(px);
            }
        }
        if (const cfg::Value* l = list_of(T, "checker")) {
            for (auto& it : l->items) {
                crt_texture t{};
                t.type = CRT_TEX_CHECKER;
                t.scale = num_of(it->get("scale"), 1.0f);
                auto sub = [&](const char* tex_key, const char* col_key) -> int32_t {
                    std::string tn = str_of(it->get(tex_key));
                    auto f2 = sf.tex_by_name.find(tn);
                    if (!tn.empty() && f2 != sf.tex_by_name.end()) return f2->second;
                    crt_texture s{};
                    s.type = CRT_TEX_SOLID; s.even = s.odd = -1;
                    color_of(it->get(col_key), s.color);
                    add("", s);
                    return (int32_t)sf.texs.size() - 1;
                };
                t.odd = sub("oddTextureName", "oddColor");
                t.even = sub("evenTextureName", "evenColor");
                add(str_of(it->get("name")), t);
            }
        }
        // textures outside the render path keep their names so materials still resolve; they are
        // sampled as mid grey and counted as unsupported
        for (const char* k : {"noise", "gradientColor", "bump", "normal", "displacement", "alpha", "emissive"}) {
            if (const cfg::Value* l = list_of(T, k)) {
                for (auto& it : l->items) {
                    crt_texture t{};
                    t.type = CRT_TEX_SOLID; t.even = t.odd = -1;
                    t.color[0] = t.color[1] = t.color[2] = 0.5f;
                    add(str_of(it->get("name")), t);
                    I.unsupported++;
                }
            }
        }
    }
    // materials
    if (const cfg::Value* M = root.get("materials")) {
        auto add = [&](const std::string& name, const crt_material& m) {
            if (!name.empty()) sf.mat_by_name[name] = (int32_t)sf.mats.size();
            sf.mats.push_back(m);
        };
        if (const cfg::Value* l = list_of(M, "lambertian")) {
            for (auto& it : l->items) {
                crt_material m{};
                m.type = CRT_MAT_LAMBERTIAN; m.texture = -1;
                std::string tn = str_of(it->get("texture"));
                if (!tn.empty()) {
                    auto f2 = sf.tex_by_name.find(tn);
                    if (f2 == sf.tex_by_name.end()) { err = "material texture '" + tn + "' not found"; return CRT_ERR_INVALID_SCENE; }
                    m.texture = f2->second;
                }
                color_of(it->get("color"), m.albedo);
                add(str_of(it->get("name")), m);
            }
        }
        if (const cfg::Value* l = list_of(M, "metal")) {
            for (auto& it : l->items) {
                crt_material m{};
                m.type = CRT_MAT_METAL; m.texture = -1;
                color_of(it->get("color"), m.albedo);
                m.fuzz = num_of(it->get("fuzziness"), 0.0f);
                add(str_of(it->get("name")), m);
            }
        }
        for (const char* k : {"glass", "dielectric"}) {
            if (const cfg::Value* l = list_of(M, k)) {
                for (auto& it : l->items) {
                    std::string name = str_of(it->get("name"));
                    if (sf.mat_by_name.count(name)) continue;   // same material under both keys
                    crt_material m{};
                    m.type = CRT_MAT_DIELECTRIC; m.texture = -1;
                    m.albedo[0] = m.albedo[1] = m.albedo[2] = 1.0f;
                    m.ior = num_of(it->get("refraction"), 1.0f);
                    add(name, m);
                }
            }
        }
        // materials outside the render path become Lambertian with their colour (mid grey if none)
        for (const char* k : {"isotropic", "anisotropic", "orennayar", "phong"}) {
            if (const cfg::Value* l = list_of(M, k)) {
                for (auto& it : l->items) {
                    crt_material m{};
                    m.type = CRT_MAT_LAMBERTIAN; m.texture = -1;
                    m.albedo[0] = m.albedo[1] = m.albedo[2] = 0.5f;
                    if (it->get("color")) color_of(it->get("color"), m.albedo);
                    add(str_of(it->get("name")), m);
                    I.unsupported++;
                }
            }
        }
    }
    auto mat_id = [&](const cfg::Value* it, int32_t& out) -> bool {
        auto f2 = sf.mat_by_name.find(str_of(it->get("material")));
        if (f2 == sf.mat_by_name.end()) return false;   // fetchMaterial fails -> primitive skipped (gpu.cu:531-534)
        out = f2->second;
        return true;
    };
    // primitives
    if (const cfg::Value* P = root.get("primitives")) {
        if (const cfg::Value* l = list_of(P, "spheres")) {
            for (auto& it : l->items) {
                if (!bool_of(it->get("active"), true)) continue;
                int32_t m;
                if (!mat_id(it.get(), m)) { I.skipped++; continue; }
                crt_sphere s{};
                vec_of(it->get("position"), s.center);
                s.radius = num_of(it->get("radius"), 1.0f);
                s.material = (uint32_t)m;
                s.mapping = crt_uvmap{1, 1, 0, 0, 1, 1};
                if (const cfg::Value* uv = it->get("uvmapping")) {
                    s.mapping.scale_u = num_of(uv->get("scale_u"), 1.0f); s.mapping.scale_v = num_of(uv->get("scale_v"), 1.0f);
                    s.mapping.offset_u = num_of(uv->get("offset_u"), 0.0f); s.mapping.offset_v = num_of(uv->get("offset_v"), 0.0f);
                    s.mapping.repeat_u = num_of(uv->get("repeat_u"), 1.0f); s.mapping.repeat_v = num_of(uv->get("repeat_v"), 1.0f);
                }
                if (it->get("transform")) I.unsupported++;
                sf.spheres.push_back(s);
            }
        }
        if (const cfg::Value* l = list_of(P, "quads")) {
            for (auto& it : l->items) {
                if (!bool_of(it->get("active"), true)) continue;
                int32_t m;
                if (!mat_id(it.get(), m)) { I.skipped++; continue; }
                crt_quad q{};
                vec_of(it->get("position"), q.corner);
                vec_of(it->get("u"), q.u);
                vec_of(it->get("v"), q.v);
                q.material = (uint32_t)m;
                if (it->get("transform")) I.unsupported++;
                sf.quads.push_back(q);
            }
        }
        if (const cfg::Value* l = list_of(P, "boxes")) {
            for (auto& it : l->items) {
                if (!bool_of(it->get("active"), true)) continue;
                int32_t m;
                if (!mat_id(it.get(), m)) { I.skipped++; continue; }
                float c[3], s[3];
                vec_of(it->get("position"), c);
                vec_of(it->get("size"), s);
                // box = six faces of [c - s/2, c + s/2] (core/src/primitives/box.cuh:41-69) as 12 triangles
                std::vector<float> pos(24);
                for (int k = 0; k < 8; ++k)
                    for (int a = 0; a < 3; ++a) pos[3 * k + a] = (k >> a & 1) ? c[a] + s[a] / 2.0f : c[a] - s[a] / 2.0f;
                sf.box_pos.push_back(pos);
                sf.box_idx.push_back({0, 1, 3, 0, 3, 2, 4, 6, 7, 4, 7, 5, 0, 4, 5, 0, 5, 1, 2, 3, 7, 2, 7, 6,
                                      0, 2, 6, 0, 6, 4, 1, 5, 7, 1, 7, 3});
                crt_mesh me{};
                me.material = (uint32_t)m;
                sf.meshes.push_back(me);
                if (it->get("transform")) I.unsupported++;
            }
        }
        for (const char* k : {"planes", "cones", "cylinders", "disks", "toruses", "volumes"})
            if (const cfg::Value* l = list_of(P, k)) I.unsupported += (uint32_t)l->items.size();
    }
    // meshes (OBJ)
    if (const cfg::Value* Ms = root.get("meshes")) {
        if (const cfg::Value* l = list_of(Ms, "obj")) {
            for (auto& it : l->items) {
                if (!bool_of(it->get("active"), true)) continue;
                int32_t m;
                if (!mat_id(it.get(), m)) { I.skipped++; continue; }
                std::string fp = resolve(sf, str_of(it->get("filepath")));
                crt_obj_mesh om{};
                if (crt_obj_load(fp.c_str(), &om) != CRT_OK) { err = "cannot load mesh " + fp; return CRT_ERR_IO; }
                if (!bool_of(it->get("use_smoothing"), true)) { std::free(om.normals); om.normals = nullptr; }
                sf.objs.push_back(om);
                crt_mesh me{};
                me.positions = om.positions; me.num_vertices = om.num_vertices;
                me.indices = om.indices; me.num_triangles = om.num_triangles;
                me.normals = om.normals; me.uvs = om.uvs; me.material = (uint32_t)m;
                sf.meshes.push_back(me);
            }
        }
    }
    // bind box geometry now that the vectors no longer move
    size_t bi = 0;
    for (auto& me : sf.meshes) {
        if (me.positions == nullptr && bi < sf.box_pos.size()) {
            me.positions = sf.box_pos[bi].data(); me.num_vertices = 8;
            me.indices = sf.box_idx[bi].data(); me.num_triangles = 12;
            ++bi;
        }
    }
    for (size_t k = 0; k < sf.texs.size(); ++k)
        if (sf.texs[k].type == CRT_TEX_IMAGE) sf.texs[k].pixels = sf.tex_pixels[k].empty() ? nullptr : sf.tex_pixels[k].data();
    crt_scene_desc& d = sf.desc;
    d.abi_version = CRT_ABI_VERSION;
    d.camera = cam;
    for (int k = 0; k < 3; ++k) d.background[k] = bg[k];
    d.materials = sf.mats.data(); d.num_materials = (uint32_t)sf.mats.size();
    d.textures = sf.texs.data(); d.num_textures = (uint32_t)sf.texs.size();
    d.lights = sf.lights.data(); d.num_lights = (uint32_t)sf.lights.size();
    d.spheres = sf.spheres.data(); d.num_spheres = (uint32_t)sf.spheres.size();
    d.quads = sf.quads.data(); d.num_quads = (uint32_t)sf.quads.size();
    d.meshes = sf.meshes.data(); d.num_meshes = (uint32_t)sf.meshes.size();
    d.instances = nullptr; d.num_instances = 0;
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
    if (rc) return set_error(rc, err);
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
