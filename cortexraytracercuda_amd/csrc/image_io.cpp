// image_io.cpp - image texture files for the `.scene` reader: PNG and JPEG.
//
// The reference loads image textures with stb_image forced to three channels
// (stbi_load(..., 3) in load_images, core/src/gpu.cu:1728-1770), any format stb reads. Here two
// decoders of this project's own cover the formats the shipped scenes use:
//   PNG   8-bit, non-interlaced, every colour type (zlib inflate, the five row filters, palettes);
//   JPEG  Huffman-coded baseline and extended sequential (SOF0 / SOF1) and progressive (SOF2)
//         frames, 8-bit, one or three components, any sampling factors, restart intervals
//         (ITU-T T.81): coefficients are decoded for the whole frame, then dequantised, inverse
//         transformed in floating point (separable 8-point IDCT), subsampled chroma interpolated
//         bilinearly between sample centres and converted from YCbCr (JFIF) to RGB.
// JPEG pixels agree with other decoders within the rounding of their IDCT and upsampling (a few
// levels at chroma edges), not bit for bit.
#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "scene_build.h"

namespace crt {

namespace {

bool read_file(const std::string& path, std::vector<uint8_t>& d, std::string& err) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) { err = "cannot open " + path; return false; }
    uint8_t buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) d.insert(d.end(), buf, buf + n);
    std::fclose(f);
    return true;
}

// ------------------------------------------------------------------ PNG
bool decode_png(const std::vector<uint8_t>& d, const std::string& path, std::vector<uint8_t>& out, uint32_t& W,
                uint32_t& H, uint32_t& C, std::string& err) {
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

// ------------------------------------------------------------------ JPEG
// zig-zag scan position -> natural (row-major) coefficient index
const uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
                             41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
                             30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct Huffman {
    bool present = false;
    uint8_t vals[256] = {};
    int maxcode[18] = {}, valptr[17] = {}, mincode[17] = {};
    uint8_t fast_len[512] = {}, fast_val[512] = {};   // codes of up to 9 bits
    bool build(const uint8_t* counts, const uint8_t* symbols, int nsym) {
        int total = 0;
        for (int l = 1; l <= 16; ++l) total += counts[l - 1];
        if (total > 256 || total > nsym) return false;
        std::memcpy(vals, symbols, (size_t)total);
        std::memset(fast_len, 0, sizeof(fast_len));
        int code = 0, k = 0;
        for (int l = 1; l <= 16; ++l) {
            valptr[l] = k;
            mincode[l] = code;
            const int n = counts[l - 1];
            for (int i = 0; i < n; ++i, ++k, ++code) {
                if (l <= 9) {
                    const int shift = 9 - l;
                    for (int s = 0; s < (1 << shift); ++s) {
                        fast_len[(code << shift) | s] = (uint8_t)l;
                        fast_val[(code << shift) | s] = vals[k];
                    }
                }
            }
            maxcode[l] = n ? code - 1 : -1;
            if (code > (1 << l)) return false;
            code <<= 1;
        }
        maxcode[17] = 0x7fffffff;
        present = true;
        return true;
    }
};

// Bit reader over an entropy-coded segment: byte stuffing (FF 00) removed; a marker stops it (zeros
// are fed from there on).
struct Bits {
    const uint8_t* p;
    const uint8_t* end;
    uint32_t buf = 0;
    int n = 0;
    bool at_marker = false;
    void fill() {
        while (n <= 24) {
            uint32_t b = 0;
            if (!at_marker && p < end) {
                if (*p == 0xFF) {
                    if (p + 1 < end && p[1] == 0x00) { b = 0xFF; p += 2; }
                    else at_marker = true;
                } else {
                    b = *p++;
                }
            }
            buf |= b << (24 - n);
            n += 8;
        }
    }
    uint32_t peek(int k) { if (n < k) fill(); return buf >> (32 - k); }
    void skip(int k) { buf <<= k; n -= k; }
    uint32_t get(int k) {
        if (k == 0) return 0;
        uint32_t v = peek(k);
        skip(k);
        return v;
    }
    int bit() { return (int)get(1); }
    void reset_at(const uint8_t* q) { p = q; buf = 0; n = 0; at_marker = false; }
    int decode(const Huffman& h) {
        const uint32_t v = peek(9);
        if (h.fast_len[v]) { skip(h.fast_len[v]); return h.fast_val[v]; }
        int code = 0, l = 0;
        do {
            code = (code << 1) | bit();
            ++l;
        } while (l <= 16 && code > h.maxcode[l]);
        if (l > 16) return 0;
        return h.vals[h.valptr[l] + code - h.mincode[l]];
    }
    int extend(int s) {
        if (s == 0) return 0;
        int v = (int)get(s);
        return v < (1 << (s - 1)) ? v - (1 << s) + 1 : v;
    }
};

struct Component {
    int id = 0, h = 1, v = 1, tq = 0;
    int bw = 0, bh = 0;            // blocks per row / column of the MCU-padded plane
    int cw = 0, ch = 0;            // blocks that cover the component's own samples
    int td = 0, ta = 0, dc_pred = 0;
    std::vector<int16_t> coef;     // 64 per block, natural order, not dequantised
};

class Jpeg {
public:
    bool decode(const std::vector<uint8_t>& d, const std::string& path, std::vector<uint8_t>& out, uint32_t& W,
                uint32_t& H, std::string& err) {
        size_t o = 2;
        bool frame = false;
        while (o + 4 <= d.size()) {
            if (d[o] != 0xFF) { ++o; continue; }
            const uint8_t m = d[o + 1];
            if (m == 0xFF) { ++o; continue; }
            if (m == 0xD9) break;                                           // EOI
            if (m == 0x01 || (m >= 0xD0 && m <= 0xD7)) { o += 2; continue; } // TEM, stray RST
            const size_t len = (size_t)d[o + 2] << 8 | d[o + 3];
            if (len < 2 || o + 2 + len > d.size()) { err = path + ": truncated JPEG segment"; return false; }
            const uint8_t* s = &d[o + 4];
            const size_t sl = len - 2;
            if (m == 0xC0 || m == 0xC1 || m == 0xC2) {
                if (!sof(s, sl, m == 0xC2, path, err)) return false;
                frame = true;
            } else if ((m >= 0xC3 && m <= 0xCF) && m != 0xC4 && m != 0xC8 && m != 0xCC) {
                err = path + ": unsupported JPEG coding (lossless, hierarchical or arithmetic)";
                return false;
            } else if (m == 0xC4) {
                if (!dht(s, sl)) { err = path + ": bad Huffman table"; return false; }
            } else if (m == 0xDB) {
                if (!dqt(s, sl)) { err = path + ": bad quantisation table"; return false; }
            } else if (m == 0xDD) {
                if (sl < 2) { err = path + ": bad restart interval"; return false; }
                restart_ = (int)(s[0] << 8 | s[1]);
            } else if (m == 0xDA) {
                if (!frame) { err = path + ": scan before frame header"; return false; }
                size_t next = 0;
                if (!scan(d, o + 2 + len, s, sl, next, path, err)) return false;
                o = next;
                continue;
            }
            o += 2 + len;
        }
        if (!frame) { err = path + ": no JPEG frame"; return false; }
        output(out);
        W = (uint32_t)w_;
        H = (uint32_t)h_;
        return true;
    }

private:
    int w_ = 0, h_ = 0, hmax_ = 1, vmax_ = 1, mcux_ = 0, mcuy_ = 0, restart_ = 0, eobrun_ = 0;
    bool progressive_ = false;
    uint16_t q_[4][64] = {};
    Huffman dc_[4], ac_[4];
    std::vector<Component> comp_;

    bool sof(const uint8_t* s, size_t n, bool prog, const std::string& path, std::string& err) {
        if (n < 6 || s[0] != 8) { err = path + ": only 8-bit JPEG is supported"; return false; }
        progressive_ = prog;
        h_ = s[1] << 8 | s[2];
        w_ = s[3] << 8 | s[4];
        const int nc = s[5];
        if (w_ <= 0 || h_ <= 0 || (uint64_t)w_ * h_ > (1ull << 28)) { err = path + ": bad JPEG size"; return false; }
        if ((nc != 1 && nc != 3) || n < 6 + 3 * (size_t)nc) { err = path + ": JPEG needs one or three components"; return false; }
        comp_.assign(nc, Component());
        for (int c = 0; c < nc; ++c) {
            comp_[c].id = s[6 + 3 * c];
            comp_[c].h = s[7 + 3 * c] >> 4;
            comp_[c].v = s[7 + 3 * c] & 15;
            comp_[c].tq = s[8 + 3 * c] & 3;
            if (comp_[c].h < 1 || comp_[c].h > 4 || comp_[c].v < 1 || comp_[c].v > 4) { err = path + ": bad sampling"; return false; }
            hmax_ = std::max(hmax_, comp_[c].h);
            vmax_ = std::max(vmax_, comp_[c].v);
        }
        mcux_ = (w_ + 8 * hmax_ - 1) / (8 * hmax_);
        mcuy_ = (h_ + 8 * vmax_ - 1) / (8 * vmax_);
        for (Component& c : comp_) {
            c.bw = mcux_ * c.h;
            c.bh = mcuy_ * c.v;
            c.cw = ((w_ * c.h + hmax_ - 1) / hmax_ + 7) / 8;
            c.ch = ((h_ * c.v + vmax_ - 1) / vmax_ + 7) / 8;
            c.coef.assign((size_t)c.bw * c.bh * 64, 0);
        }
        return true;
    }
    bool dht(const uint8_t* s, size_t n) {
        size_t o = 0;
        while (o + 17 <= n) {
            const int tc = s[o] >> 4, th = s[o] & 15;
            if (tc > 1 || th > 3) return false;
            int total = 0;
            for (int k = 0; k < 16; ++k) total += s[o + 1 + k];
            if (o + 17 + (size_t)total > n) return false;
            Huffman& h = tc ? ac_[th] : dc_[th];
            if (!h.build(s + o + 1, s + o + 17, total)) return false;
            o += 17 + total;
        }
        return o == n;
    }
    bool dqt(const uint8_t* s, size_t n) {
        size_t o = 0;
        while (o < n) {
            const int pq = s[o] >> 4, tq = s[o] & 15;
            if (tq > 3 || pq > 1) return false;
            const size_t need = 1 + 64 * (pq + 1);
            if (o + need > n) return false;
            for (int k = 0; k < 64; ++k)
                q_[tq][kZigzag[k]] = pq ? (uint16_t)(s[o + 1 + 2 * k] << 8 | s[o + 2 + 2 * k]) : s[o + 1 + k];
            o += need;
        }
        return true;
    }

    // one block of a scan
    void block(Bits& b, Component& c, int16_t* blk, int ss, int se, int ah, int al) {
        if (!progressive_) {
            const int t = b.decode(dc_[c.td]);
            c.dc_pred += b.extend(t);
            blk[0] = (int16_t)c.dc_pred;
            for (int k = 1; k < 64;) {
                const int rs = b.decode(ac_[c.ta]);
                const int r = rs >> 4, s = rs & 15;
                if (s == 0) {
                    if (r != 15) break;
                    k += 16;
                    continue;
                }
                k += r;
                if (k > 63) break;
                blk[kZigzag[k]] = (int16_t)b.extend(s);
                ++k;
            }
            return;
        }
        if (ss == 0) {   // DC scan of a progressive frame
            if (ah == 0) {
                const int t = b.decode(dc_[c.td]);
                c.dc_pred += b.extend(t);
                blk[0] = (int16_t)(c.dc_pred * (1 << al));
            } else if (b.bit()) {
                blk[0] = (int16_t)(blk[0] | (1 << al));
            }
            return;
        }
        if (ah == 0) {   // AC first pass
            if (eobrun_ > 0) { --eobrun_; return; }
            for (int k = ss; k <= se;) {
                const int rs = b.decode(ac_[c.ta]);
                const int r = rs >> 4, s = rs & 15;
                if (s == 0) {
                    if (r < 15) {
                        eobrun_ = (1 << r) - 1;
                        if (r) eobrun_ += (int)b.get(r);
                        break;
                    }
                    k += 16;
                    continue;
                }
                k += r;
                if (k > 63) break;
                blk[kZigzag[k]] = (int16_t)(b.extend(s) * (1 << al));
                ++k;
            }
            return;
        }
        // AC refinement: one correction bit for every coefficient already nonzero that is passed
        const int p1 = 1 << al, m1 = -(1 << al);
        int k = ss;
        auto refine = [&](int16_t& v) {
            if (b.bit() && (v & p1) == 0) v = (int16_t)(v + (v >= 0 ? p1 : m1));
        };
        if (eobrun_ == 0) {
            for (; k <= se; ++k) {
                const int rs = b.decode(ac_[c.ta]);
                int r = rs >> 4;
                const int s = rs & 15;
                int val = 0;
                if (s) {
                    val = b.bit() ? p1 : m1;
                } else if (r != 15) {
                    eobrun_ = 1 << r;
                    if (r) eobrun_ += (int)b.get(r);
                    break;
                }
                while (k <= se) {
                    int16_t& v = blk[kZigzag[k]];
                    if (v != 0) {
                        refine(v);
                    } else {
                        if (r == 0) {
                            if (val) v = (int16_t)val;
                            break;
                        }
                        --r;
                    }
                    ++k;
                }
            }
        }
        if (eobrun_ > 0) {
            for (; k <= se; ++k) {
                int16_t& v = blk[kZigzag[k]];
                if (v != 0) refine(v);
            }
            --eobrun_;
        }
    }

    bool scan(const std::vector<uint8_t>& d, size_t data, const uint8_t* s, size_t n, size_t& next, const std::string& path,
              std::string& err) {
        if (n < 1) { err = path + ": bad scan header"; return false; }
        const int ns = s[0];
        if (ns < 1 || ns > 4 || n < 4 + 2 * (size_t)ns) { err = path + ": bad scan header"; return false; }
        std::vector<int> ci;
        for (int k = 0; k < ns; ++k) {
            int found = -1;
            for (size_t c = 0; c < comp_.size(); ++c) if (comp_[c].id == s[1 + 2 * k]) found = (int)c;
            if (found < 0) { err = path + ": scan names an unknown component"; return false; }
            comp_[found].td = s[2 + 2 * k] >> 4;
            comp_[found].ta = s[2 + 2 * k] & 15;
            if (comp_[found].td > 3 || comp_[found].ta > 3) { err = path + ": bad table selector"; return false; }
            ci.push_back(found);
        }
        const int ss = s[1 + 2 * ns], se = s[2 + 2 * ns], ah = s[3 + 2 * ns] >> 4, al = s[3 + 2 * ns] & 15;
        if (!progressive_) { if (ss != 0 || se != 63) { err = path + ": bad sequential scan"; return false; } }
        else if (ss > se || se > 63 || (ss == 0 && se != 0) || (ss > 0 && ns != 1) || al > 13) {
            err = path + ": bad progressive scan";
            return false;
        }
        for (int c : ci) {
            const bool need_dc = !progressive_ || (ss == 0 && ah == 0);
            const bool need_ac = !progressive_ || ss > 0;
            if ((need_dc && !dc_[comp_[c].td].present) || (need_ac && !ac_[comp_[c].ta].present)) {
                err = path + ": missing Huffman table";
                return false;
            }
        }
        Bits b{d.data() + data, d.data() + d.size()};
        for (int c : ci) comp_[c].dc_pred = 0;
        eobrun_ = 0;
        int left = restart_;
        auto restart = [&]() {
            // byte-align, step over the RSTn marker, reset the predictors and the EOB run
            const uint8_t* q = b.p;
            while (q + 1 < b.end && !(q[0] == 0xFF && q[1] >= 0xD0 && q[1] <= 0xD7)) {
                if (q[0] == 0xFF && q[1] != 0x00) break;
                ++q;
            }
            if (q + 1 < b.end && q[0] == 0xFF && q[1] >= 0xD0 && q[1] <= 0xD7) q += 2;
            b.reset_at(q);
            for (int c : ci) comp_[c].dc_pred = 0;
            eobrun_ = 0;
            left = restart_;
        };
        if (ns == 1) {
            Component& c = comp_[ci[0]];
            for (int by = 0; by < c.ch; ++by)
                for (int bx = 0; bx < c.cw; ++bx) {
                    if (restart_ && left == 0) restart();
                    block(b, c, &c.coef[((size_t)by * c.bw + bx) * 64], ss, se, ah, al);
                    if (restart_) --left;
                }
        } else {
            for (int my = 0; my < mcuy_; ++my)
                for (int mx = 0; mx < mcux_; ++mx) {
                    if (restart_ && left == 0) restart();
                    for (int k : ci) {
                        Component& c = comp_[k];
                        for (int v = 0; v < c.v; ++v)
                            for (int h = 0; h < c.h; ++h)
                                block(b, c, &c.coef[((size_t)(my * c.v + v) * c.bw + mx * c.h + h) * 64], ss, se, ah, al);
                    }
                    if (restart_) --left;
                }
        }
        // the next marker after the entropy-coded data (restart markers belong to the scan)
        const uint8_t* q = b.p;
        while (q + 1 < b.end && !(q[0] == 0xFF && q[1] != 0x00 && !(q[1] >= 0xD0 && q[1] <= 0xD7))) ++q;
        next = (size_t)(q - d.data());
        return true;
    }

    void output(std::vector<uint8_t>& out) {
        // separable IDCT: f(x) = sum_u C(u)/2 F(u) cos((2x + 1) u pi / 16), C(0) = 1/sqrt(2)
        float t[8][8];
        for (int x = 0; x < 8; ++x)
            for (int u = 0; u < 8; ++u)
                t[x][u] = (u == 0 ? (float)(1.0 / std::sqrt(2.0)) : 1.0f) * 0.5f *
                          (float)std::cos((2 * x + 1) * u * 3.14159265358979323846 / 16.0);
        std::vector<std::vector<uint8_t>> planes(comp_.size());
        for (size_t ci = 0; ci < comp_.size(); ++ci) {
            Component& c = comp_[ci];
            const int pw = c.bw * 8;
            planes[ci].assign((size_t)pw * c.bh * 8, 0);
            const uint16_t* q = q_[c.tq];
            for (int by = 0; by < c.bh; ++by)
                for (int bx = 0; bx < c.bw; ++bx) {
                    const int16_t* blk = &c.coef[((size_t)by * c.bw + bx) * 64];
                    float f[64], r[64];
                    for (int k = 0; k < 64; ++k) f[k] = (float)blk[k] * (float)q[k];
                    for (int y = 0; y < 8; ++y)          // rows: u -> x
                        for (int x = 0; x < 8; ++x) {
                            float a = 0.0f;
                            for (int u = 0; u < 8; ++u) a += t[x][u] * f[y * 8 + u];
                            r[y * 8 + x] = a;
                        }
                    for (int x = 0; x < 8; ++x)          // columns: v -> y
                        for (int y = 0; y < 8; ++y) {
                            float a = 0.0f;
                            for (int v = 0; v < 8; ++v) a += t[y][v] * r[v * 8 + x];
                            const int px = (int)std::lround(a + 128.0f);
                            planes[ci][(size_t)(by * 8 + y) * pw + bx * 8 + x] = (uint8_t)(px < 0 ? 0 : (px > 255 ? 255 : px));
                        }
                }
        }
        // subsampled components: bilinear between the centres of the component's samples (sample i
        // of a component with factor h covers full-resolution x in [i, i + 1) * hmax / h), edges clamped
        out.resize((size_t)w_ * h_ * 3);
        auto sample = [&](size_t ci, int x, int y) -> float {
            const Component& c = comp_[ci];
            const int pw = c.bw * 8;
            const std::vector<uint8_t>& P = planes[ci];
            if (c.h == hmax_ && c.v == vmax_) return P[(size_t)y * pw + x];
            const int sw = (w_ * c.h + hmax_ - 1) / hmax_, sh = (h_ * c.v + vmax_ - 1) / vmax_;
            const float fx = ((float)x + 0.5f) * (float)c.h / (float)hmax_ - 0.5f;
            const float fy = ((float)y + 0.5f) * (float)c.v / (float)vmax_ - 0.5f;
            const int x0 = (int)std::floor(fx), y0 = (int)std::floor(fy);
            const float ax = fx - (float)x0, ay = fy - (float)y0;
            auto at = [&](int xx, int yy) -> float {
                xx = xx < 0 ? 0 : (xx >= sw ? sw - 1 : xx);
                yy = yy < 0 ? 0 : (yy >= sh ? sh - 1 : yy);
                return P[(size_t)yy * pw + xx];
            };
            const float top = at(x0, y0) * (1.0f - ax) + at(x0 + 1, y0) * ax;
            const float bot = at(x0, y0 + 1) * (1.0f - ax) + at(x0 + 1, y0 + 1) * ax;
            return top * (1.0f - ay) + bot * ay;
        };
        for (int y = 0; y < h_; ++y)
            for (int x = 0; x < w_; ++x) {
                float s[3] = {0, 0, 0};
                for (size_t ci = 0; ci < comp_.size(); ++ci) s[ci] = sample(ci, x, y);
                uint8_t* o = &out[((size_t)y * w_ + x) * 3];
                if (comp_.size() == 1) { o[0] = o[1] = o[2] = (uint8_t)s[0]; continue; }
                const float Y = s[0], cb = s[1] - 128.0f, cr = s[2] - 128.0f;
                const float rgb[3] = {Y + 1.402f * cr, Y - 0.344136f * cb - 0.714136f * cr, Y + 1.772f * cb};
                for (int k = 0; k < 3; ++k) {
                    const int v = (int)std::lround(rgb[k]);
                    o[k] = (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
                }
            }
    }
};

}  // namespace

bool read_image(const std::string& path, std::vector<uint8_t>& out, uint32_t& W, uint32_t& H, uint32_t& C, std::string& err) {
    std::vector<uint8_t> d;
    if (!read_file(path, d, err)) return false;
    static const uint8_t png_sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    if (d.size() >= 8 && std::memcmp(d.data(), png_sig, 8) == 0) return decode_png(d, path, out, W, H, C, err);
    if (d.size() >= 4 && d[0] == 0xFF && d[1] == 0xD8) {
        Jpeg j;
        C = 3;
        return j.decode(d, path, out, W, H, err);
    }
    err = path + ": not a PNG or JPEG file";
    return false;
}

}  // namespace crt
