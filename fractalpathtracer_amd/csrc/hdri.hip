// HDRI environment LUT: Radiance RGBE (.hdr) reader, box downsampling to the
// config's 32 x 16 FP16 table, and the config hydration that every render
// entry point performs (metal_bridge.mm:2163-2300, called from
// metal_bridge.mm:4331 and 5880).  Host code only.
//
// Format handled (as the reference): header lines up to an empty line, of
// which one must be FORMAT=32-bit_rle_rgbe; a resolution line "<s>Y h <s>X w"
// with 8 <= w <= 32767; every scanline in the adaptive run-length form
// (2, 2, w >> 8, w & 255, then each of the four channels as packets: a count
// above 128 repeats the next byte count - 128 times, otherwise count literal
// bytes follow).  A pixel (r, g, b, e) decodes to (c + 0.5) 2^(e - 136), or 0
// when e = 0.  Rows are stored top-down when the Y sign is '-', columns
// left-right when the X sign is '+'.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "fpt_internal.h"

namespace fpt {

namespace {

struct RgbImage {
    uint32_t width = 0, height = 0;
    std::vector<float> rgb;  // row-major, top row first
};

bool fail(char *error, size_t error_len, const char *msg, const char *path) {
    if (error && error_len) std::snprintf(error, error_len, msg, path);
    return false;
}

bool read_rgbe(const char *path, RgbImage &img, char *error, size_t error_len) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return fail(error, error_len, "failed to open HDRI %s", path);
    std::string line;
    bool rle_rgbe = false;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.compare(0, 22, "FORMAT=32-bit_rle_rgbe") == 0) rle_rgbe = true;
        if (line.empty()) break;
    }
    if (!std::getline(in, line)) return fail(error, error_len, "missing HDRI resolution in %s", path);
    char ys = 0, xs = 0;
    int h = 0, w = 0;
    if (std::sscanf(line.c_str(), " %cY %d %cX %d", &ys, &h, &xs, &w) != 4 || w <= 0 || h <= 0) {
        if (error && error_len) std::snprintf(error, error_len, "unsupported HDRI resolution line '%s'", line.c_str());
        return false;
    }
    if (!rle_rgbe || w < 8 || w > 32767) return fail(error, error_len, "unsupported non-RLE Radiance HDRI %s", path);
    img.width = uint32_t(w);
    img.height = uint32_t(h);
    img.rgb.assign(size_t(w) * size_t(h) * 3u, 0.0f);
    std::vector<uint8_t> row(size_t(w) * 4u);  // channel-interleaved r g b e
    for (int fy = 0; fy < h; ++fy) {
        unsigned char hdr[4] = {};
        in.read(reinterpret_cast<char *>(hdr), 4);
        if (!in || hdr[0] != 2u || hdr[1] != 2u || ((uint32_t(hdr[2]) << 8) | hdr[3]) != uint32_t(w)) {
            if (error && error_len) std::snprintf(error, error_len, "invalid HDRI scanline %d in %s", fy, path);
            return false;
        }
        for (int c = 0; c < 4; ++c) {
            int x = 0;
            while (x < w) {
                int code = in.get();
                if (code == EOF || code == 0) {
                    if (error && error_len) std::snprintf(error, error_len, "truncated HDRI scanline %d in %s", fy, path);
                    return false;
                }
                if (code > 128) {  // run
                    const int n = code - 128;
                    const int v = in.get();
                    if (v == EOF || x + n > w) return fail(error, error_len, "corrupt HDRI run in %s", path);
                    for (int i = 0; i < n; ++i) row[size_t(x++) * 4u + c] = uint8_t(v);
                } else {  // literal bytes
                    if (x + code > w) return fail(error, error_len, "corrupt HDRI literal in %s", path);
                    for (int i = 0; i < code; ++i) {
                        const int v = in.get();
                        if (v == EOF) return fail(error, error_len, "truncated HDRI literal in %s", path);
                        row[size_t(x++) * 4u + c] = uint8_t(v);
                    }
                }
            }
        }
        const int oy = ys == '-' ? fy : h - 1 - fy;
        for (int fx = 0; fx < w; ++fx) {
            const int ox = xs == '+' ? fx : w - 1 - fx;
            const uint8_t *px = &row[size_t(fx) * 4u];
            const float scale = px[3] == 0u ? 0.0f : std::ldexp(1.0f, int(px[3]) - 136);
            float *o = &img.rgb[(size_t(oy) * size_t(w) + size_t(ox)) * 3u];
            for (int c = 0; c < 3; ++c) o[c] = (float(px[c]) + 0.5f) * scale;
        }
    }
    return true;
}

}  // namespace

// Half-precision bits of v with the reference's rounding (metal_bridge.mm:
// 2163-2176): add half a unit of the kept precision to the discarded bits and
// truncate (ties away from zero), subnormals by shifting in the implicit bit,
// magnitudes below 2^-25 to signed zero, overflow, infinity and NaN to
// infinity.
uint16_t half_bits_round_up(float v) {
    uint32_t b;
    std::memcpy(&b, &v, 4);
    const uint32_t sign = (b >> 16) & 0x8000u;
    const int e = int((b >> 23) & 0xffu) - 112;  // rebias 127 -> 15
    uint32_t m = b & 0x7fffffu;
    if (e >= 31) return uint16_t(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return uint16_t(sign);
        m = (m | 0x800000u) >> uint32_t(1 - e);
        return uint16_t(sign | ((m + 0x1000u) >> 13));
    }
    return uint16_t(sign | ((uint32_t(e) << 10) + ((m + 0x1000u) >> 13)));
}

bool hydrate_hdri(FptRenderConfig &cfg, char *error, size_t error_len) {
    cfg.hdri_width = 0u;
    cfg.hdri_height = 0u;
    std::memset(cfg.hdri_lut, 0, sizeof(cfg.hdri_lut));
    if (cfg.hdri_enabled == 0u || cfg.world[0] != 2.0f) return true;
    char path[FPT_HDRI_PATH_CAPACITY + 1];
    std::memcpy(path, cfg.hdri_path, FPT_HDRI_PATH_CAPACITY);
    path[FPT_HDRI_PATH_CAPACITY] = 0;
    if (!path[0]) {
        if (error && error_len) std::snprintf(error, error_len, "HDRI world mode requires world.hdri");
        return false;
    }
    RgbImage img;
    if (!read_rgbe(path, img, error, error_len)) return false;
    const uint32_t W = FPT_HDRI_LUT_WIDTH, H = FPT_HDRI_LUT_HEIGHT;
    cfg.hdri_width = W;
    cfg.hdri_height = H;
    // Each LUT texel is the mean of its source box [x w / W, (x + 1) w / W)
    // x [y h / H, (y + 1) h / H), at least one pixel wide, summed row by row.
    for (uint32_t ty = 0; ty < H; ++ty) {
        const uint32_t y0 = uint32_t(uint64_t(ty) * img.height / H);
        const uint32_t y1 = std::min(std::max(y0 + 1u, uint32_t(uint64_t(ty + 1u) * img.height / H)), img.height);
        for (uint32_t tx = 0; tx < W; ++tx) {
            const uint32_t x0 = uint32_t(uint64_t(tx) * img.width / W);
            const uint32_t x1 = std::min(std::max(x0 + 1u, uint32_t(uint64_t(tx + 1u) * img.width / W)), img.width);
            float acc[3] = {0.f, 0.f, 0.f};
            uint32_t n = 0;
            for (uint32_t sy = y0; sy < y1; ++sy)
                for (uint32_t sx = x0; sx < x1; ++sx, ++n)
                    for (int c = 0; c < 3; ++c) acc[c] += img.rgb[(size_t(sy) * img.width + sx) * 3u + c];
            for (int c = 0; c < 3; ++c)
                cfg.hdri_lut[(size_t(ty) * W + tx) * 3u + c] = half_bits_round_up(acc[c] / float(std::max(n, 1u)));
        }
    }
    return true;
}

}  // namespace fpt
