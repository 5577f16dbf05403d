// crt_cli.cpp - command-line renderer with the reference executable's interface.
//
//   crt_render -scene file.scene -width W -height H -ratio a:b -spp S -maxdepth D -save out.png
//              [-quiet] [-mode N] [-rng pcg|minstd] [-seed N] [-noprint]
//
// Mirrors core/src/main.cpp:10-42 and render_parameters::getArgs (core/src/misc/render_parameters.h:
// 23-105): the same flags and defaults (width 256, ratio 16:9, height = width / ratio, spp 100,
// maxdepth 100), the command line overriding the scene file's width, spp, depth and aspect ratio
// (core/src/scenes/scene_manager.cpp:82-86), the PNG written with rows top-down, and the per-pixel
// stdout protocol `p x y r g b` followed by four black padding rows (core/src/gpu.cu:801,
// 2088-2094) that the reference UI parses (UI/Examples/UI/main.cpp:241-246). `-mode` is accepted
// and ignored as in the reference. The render itself goes through the C-ABI.
#include <zlib.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cortex_rt.h"

static bool write_png(const char* path, const uint8_t* rgb, uint32_t W, uint32_t H) {
    std::vector<uint8_t> raw((size_t)(W * 3 + 1) * H);
    for (uint32_t y = 0; y < H; ++y) {
        raw[(size_t)y * (W * 3 + 1)] = 0;
        std::memcpy(&raw[(size_t)y * (W * 3 + 1) + 1], rgb + (size_t)y * W * 3, (size_t)W * 3);
    }
    uLongf zl = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zl);
    if (compress2(z.data(), &zl, raw.data(), (uLong)raw.size(), 6) != Z_OK) return false;
    FILE* f = std::fopen(path, "wb");
    if (!f) return false;
    auto be = [](uint8_t* p, uint32_t v) { p[0] = v >> 24; p[1] = v >> 16; p[2] = v >> 8; p[3] = v; };
    auto chunk = [&](const char* tag, const uint8_t* data, uint32_t len) {
        uint8_t hdr[8];
        be(hdr, len);
        std::memcpy(hdr + 4, tag, 4);
        std::fwrite(hdr, 1, 8, f);
        if (len) std::fwrite(data, 1, len, f);
        uLong crc = crc32(0L, (const Bytef*)tag, 4);
        if (len) crc = crc32(crc, data, len);
        uint8_t c[4];
        be(c, (uint32_t)crc);
        std::fwrite(c, 1, 4, f);
    };
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::fwrite(sig, 1, 8, f);
    uint8_t ihdr[13];
    be(ihdr, W); be(ihdr + 4, H);
    ihdr[8] = 8; ihdr[9] = 2; ihdr[10] = 0; ihdr[11] = 0; ihdr[12] = 0;
    chunk("IHDR", ihdr, 13);
    chunk("IDAT", z.data(), (uint32_t)zl);
    chunk("IEND", nullptr, 0);
    return std::fclose(f) == 0;
}

int main(int argc, char** argv) {
    // render_parameters defaults (core/src/misc/render_parameters.h:13-17)
    bool quiet = false, print_pixels = true;
    float ratio = 16.0f / 9.0f;
    int width = 256;
    int height = (int)(width / ratio);
    int spp = 100, depth = 100;
    std::string scene, save = "output.png", rng = "minstd";
    unsigned long long seed = 2580;
    for (int k = 1; k < argc; ++k) {
        std::string a = argv[k];
        if (a.empty() || a[0] != '-') continue;
        std::string p = a.substr(1);
        std::string v = k + 1 < argc ? argv[k + 1] : "";
        if (p == "quiet") quiet = true;
        else if (p == "noprint") print_pixels = false;
        else if (p == "width" && !v.empty()) width = (int)std::strtoul(v.c_str(), nullptr, 10);
        else if (p == "height" && !v.empty()) height = (int)std::strtoul(v.c_str(), nullptr, 10);
        else if (p == "ratio" && !v.empty()) {
            double p1 = 0, p2 = 0;
            if (std::sscanf(v.c_str(), "%lf:%lf", &p1, &p2) == 2 && p1 > 0 && p2 > 0) {
                ratio = (float)(p1 / p2);
                if (ratio > 1) height = (int)(width / ratio); else width = (int)(height * ratio);
            }
        } else if (p == "spp" && !v.empty()) spp = (int)std::strtoul(v.c_str(), nullptr, 10);
        else if (p == "maxdepth" && !v.empty()) depth = (int)std::strtoul(v.c_str(), nullptr, 10);
        else if (p == "scene" && !v.empty()) scene = v;
        else if (p == "save" && !v.empty()) save = v;
        else if (p == "rng" && !v.empty()) rng = v;
        else if (p == "seed" && !v.empty()) seed = std::strtoull(v.c_str(), nullptr, 10);
    }
    if (scene.empty()) { std::fprintf(stderr, "[ERROR] no -scene given\n"); return 1; }
    std::printf("[INFO] Ready !\n");
    crt_scenefile* sf = nullptr;
    if (crt_scenefile_load(scene.c_str(), &sf) != CRT_OK) {
        std::fprintf(stderr, "[ERROR] %s\n", crt_last_error());
        return 1;
    }
    crt_scene_desc desc;
    crt_scenefile_info info;
    crt_scenefile_desc(sf, &desc, &info);
    desc.camera.aspect_ratio = ratio;   // command line is stronger than the .scene (scene_manager.cpp:82-86)
    if (info.unsupported && !quiet)
        std::fprintf(stderr, "[WARNING] %u scene entries outside the render path were skipped\n", info.unsupported);
    auto t0 = std::chrono::steady_clock::now();
    crt_context* ctx = nullptr;
    crt_config cc{};
    if (crt_create(&cc, &ctx) != CRT_OK) {
        std::printf("[ERROR] %s\n", crt_last_error());   // reference: no GPU -> message and return
        crt_scenefile_close(sf);
        return 0;
    }
    crt_scene_id sid;
    if (crt_scene_upload(ctx, &desc, &sid) != CRT_OK) {
        std::fprintf(stderr, "[ERROR] %s\n", crt_last_error());
        return 99;
    }
    crt_render_desc rd{};
    rd.width = (uint32_t)width; rd.height = (uint32_t)height; rd.spp = (uint32_t)spp; rd.max_depth = (uint32_t)depth;
    rd.seed = seed;
    rd.rng_mode = rng == "pcg" ? CRT_RNG_PCG : CRT_RNG_MINSTD;
    std::printf("[INFO] Rendering %dx%d %d samples > %s\n", width, height, spp, save.c_str());
    std::vector<uint8_t> rgb((size_t)width * height * 3);
    crt_image img{rgb.data(), nullptr};
    crt_stats st{};
    if (crt_render(ctx, sid, &rd, &img, &st) != CRT_OK) {
        std::fprintf(stderr, "[ERROR] %s\n", crt_last_error());
        return 99;
    }
    if (print_pixels) {
        for (int y = 0; y < height; ++y)
            for (int x = 0; x < width; ++x) {
                const uint8_t* px = &rgb[3 * ((size_t)y * width + x)];
                std::printf("p %d %d %u %u %u\n", x, y, px[0], px[1], px[2]);
            }
        for (int jj = 0; jj < 4; ++jj)
            for (int ii = 0; ii < width; ++ii) std::printf("p %d %d 0 0 0\n", ii, height - jj - 1);
    }
    if (!write_png(save.c_str(), rgb.data(), (uint32_t)width, (uint32_t)height)) {
        std::fprintf(stderr, "[ERROR] cannot write %s\n", save.c_str());
        return 1;
    }
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!quiet)
        std::printf("[INFO] %llu rays in %.3f ms of GPU time (%.1f Mrays/s); total %.3f s\n",
                    (unsigned long long)st.rays, st.gpu_ms, st.rays / (st.gpu_ms * 1e3), secs);
    std::printf("[INFO] Finished !\n");
    crt_scene_release(ctx, sid);
    crt_destroy(ctx);
    crt_scenefile_close(sf);
    return 0;
}
