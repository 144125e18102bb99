// C-ABI layer of libfpt_hip.so: the reference's entry points
// (src/metal_bridge.h:404-486, behaviour of src/metal_bridge.mm) and the new
// fpt_hip_* entry points declared in include/fpt_abi.h.
//
// Every call is synchronous on its own HIP stream, allocates what it needs
// and frees it before returning; the caller owns every pointer
// (metal_bridge.mm:4298-5863 ownership rules).
#include <hip/hip_runtime.h>
#include <zlib.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fpt_abi.h"
#include "fpt_internal.h"
#include "learned.h"
#include "specialize.h"
#include "voxel_api.h"

namespace {

thread_local std::string g_last_error;

void set_error(char *error, size_t error_len, const char *fmt, ...) {
    char buf[1024];
    va_list args;
    va_start(args, fmt);
    vsnprintf(buf, sizeof(buf), fmt, args);
    va_end(args);
    g_last_error = buf;
    if (error && error_len) {
        std::snprintf(error, error_len, "%s", buf);
    }
}

#define HIP_TRY(expr, what)                                                                     \
    do {                                                                                        \
        hipError_t _e = (expr);                                                                 \
        if (_e != hipSuccess) {                                                                 \
            set_error(error, error_len, "%s failed: %s", what, hipGetErrorString(_e));          \
            return 1;                                                                           \
        }                                                                                       \
    } while (0)

// Device buffer owned for the duration of one call.
struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t n) { return hipMalloc(&p, n ? n : 1); }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

struct Stream {
    hipStream_t s = nullptr;
    ~Stream() {
        if (s) (void)hipStreamDestroy(s);
    }
};

int check_device(char *error, size_t error_len) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        set_error(error, error_len, "no HIP device is available (%s)", hipGetErrorString(e));
        return 1;
    }
    return 0;
}

uint32_t clamp_samples(uint32_t s) { return std::min<uint32_t>(std::max<uint32_t>(s, 1u), 512u); }

// Reference accumulation policy (metal_bridge.mm:4531-4564): all samples in
// one dispatch when spp <= 16, else chunks of sdf_chunk_samples (default 8).
// The running mean makes every policy produce the same image.
uint32_t chunk_for(const FptRenderConfig &cfg) {
    const uint32_t spp = clamp_samples(cfg.samples);
    if (cfg.sdf_accumulation_mode == FPT_SDF_ACCUMULATION_PER_SAMPLE) return 1u;
    if (cfg.sdf_accumulation_mode == FPT_SDF_ACCUMULATION_BATCH) return spp;
    if (cfg.sdf_accumulation_mode == FPT_SDF_ACCUMULATION_AUTO && spp <= 16u) return spp;
    uint32_t c = cfg.sdf_chunk_samples ? cfg.sdf_chunk_samples : 8u;
    return std::min<uint32_t>(std::max<uint32_t>(c, 1u), 64u);
}

int validate_config(const FptRenderConfig *config, char *error, size_t error_len) {
    if (!config) {
        set_error(error, error_len, "missing render config");
        return 1;
    }
    if (config->width == 0 || config->height == 0) {
        set_error(error, error_len, "invalid render size %ux%u", config->width, config->height);
        return 1;
    }
    if (config->width > 16384 || config->height > 16384) {
        set_error(error, error_len, "render size %ux%u exceeds 16384x16384", config->width, config->height);
        return 1;
    }
    if (config->renderer_backend != FPT_RENDERER_SDF && config->renderer_backend != FPT_RENDERER_VOXEL) {
        set_error(error, error_len, "renderer backend %u is not supported by this library", config->renderer_backend);
        return 1;
    }
    if (config->sdf_id == FPT_SDF_MANDELBULBER && config->renderer_backend != FPT_RENDERER_SDF) {
        set_error(error, error_len, "imported Mandelbulber scenes render on the SDF backend only");
        return 1;
    }
    return 0;
}

// Imported Mandelbulber scenes (the native KIFS field): batches, learned
// stepping, the persistent-kernel autotuner and the voxel builder refuse
// them.
int reject_mandelbulber(const FptRenderConfig *config, const char *what, char *error, size_t error_len) {
    if (!config || config->sdf_id != FPT_SDF_MANDELBULBER) return 0;
    set_error(error, error_len, "%s does not support imported Mandelbulber scenes", what);
    return 1;
}

// ---------------------------------------------------------------------------
// PNG writer (RGBA8, zlib).
// ---------------------------------------------------------------------------
void put_be32(std::vector<uint8_t> &v, uint32_t x) {
    v.push_back(uint8_t(x >> 24)); v.push_back(uint8_t(x >> 16)); v.push_back(uint8_t(x >> 8)); v.push_back(uint8_t(x));
}
void put_chunk(std::vector<uint8_t> &out, const char *type, const std::vector<uint8_t> &data) {
    put_be32(out, uint32_t(data.size()));
    size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    uLong crc = crc32(0L, out.data() + start, uInt(out.size() - start));
    put_be32(out, uint32_t(crc));
}

bool write_png(const char *path, uint32_t w, uint32_t h, const std::vector<uint8_t> &rgba, char *error,
               size_t error_len) {
    std::vector<uint8_t> raw;
    raw.reserve(size_t(h) * (size_t(w) * 4 + 1));
    for (uint32_t y = 0; y < h; ++y) {
        raw.push_back(0);
        raw.insert(raw.end(), rgba.begin() + size_t(y) * w * 4, rgba.begin() + size_t(y + 1) * w * 4);
    }
    uLongf zlen = compressBound(uLong(raw.size()));
    std::vector<uint8_t> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), uLong(raw.size()), 6) != Z_OK) {
        set_error(error, error_len, "PNG compression failed");
        return false;
    }
    z.resize(zlen);
    std::vector<uint8_t> png = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, w);
    put_be32(ihdr, h);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});
    put_chunk(png, "IHDR", ihdr);
    put_chunk(png, "IDAT", z);
    put_chunk(png, "IEND", {});
    FILE *f = std::fopen(path, "wb");
    if (!f) {
        set_error(error, error_len, "cannot open %s for writing", path);
        return false;
    }
    size_t n = std::fwrite(png.data(), 1, png.size(), f);
    std::fclose(f);
    if (n != png.size()) {
        set_error(error, error_len, "short write to %s", path);
        return false;
    }
    return true;
}

struct Image {
    uint32_t width = 0, height = 0;
    std::vector<uint8_t> rgba;
};

uint32_t get_be32(const uint8_t *p) { return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | p[3]; }

// PNG decoder for 8-bit, non-interlaced grey / grey+alpha / RGB / RGBA
// images (all five row filters), expanded to RGBA.
bool read_png(const char *path, Image &img, char *error, size_t error_len) {
    FILE *f = path ? std::fopen(path, "rb") : nullptr;
    if (!f) {
        set_error(error, error_len, "failed to read image %s", path ? path : "(null)");
        return false;
    }
    std::vector<uint8_t> data;
    uint8_t buf[65536];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) data.insert(data.end(), buf, buf + n);
    std::fclose(f);
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    if (data.size() < 8 || std::memcmp(data.data(), sig, 8) != 0) {
        set_error(error, error_len, "%s is not a PNG image", path);
        return false;
    }
    uint32_t w = 0, h = 0;
    int channels = 0;
    std::vector<uint8_t> z;
    for (size_t pos = 8; pos + 12 <= data.size();) {
        const uint32_t len = get_be32(&data[pos]);
        if (pos + 12 + size_t(len) > data.size()) break;
        const uint8_t *type = &data[pos + 4], *body = &data[pos + 8];
        if (!std::memcmp(type, "IHDR", 4) && len >= 13) {
            w = get_be32(body);
            h = get_be32(body + 4);
            const uint8_t depth = body[8], ctype = body[9], interlace = body[12];
            channels = ctype == 0 ? 1 : ctype == 4 ? 2 : ctype == 2 ? 3 : ctype == 6 ? 4 : 0;
            if (depth != 8 || channels == 0 || interlace != 0) {
                set_error(error, error_len, "%s: only 8-bit non-interlaced grey/RGB(A) PNGs are supported", path);
                return false;
            }
        } else if (!std::memcmp(type, "IDAT", 4)) {
            z.insert(z.end(), body, body + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            break;
        }
        pos += 12 + size_t(len);
    }
    if (w == 0 || h == 0 || w > 65536 || h > 65536) {
        set_error(error, error_len, "%s: missing or invalid PNG header", path);
        return false;
    }
    const size_t stride = size_t(w) * channels;
    std::vector<uint8_t> raw(size_t(h) * (stride + 1));
    z_stream zs{};
    zs.next_in = z.data();
    zs.avail_in = uInt(z.size());
    zs.next_out = raw.data();
    zs.avail_out = uInt(raw.size());
    int zr = inflateInit(&zs);
    if (zr == Z_OK) zr = inflate(&zs, Z_FINISH);
    inflateEnd(&zs);
    if (zr != Z_STREAM_END || zs.total_out != raw.size()) {
        set_error(error, error_len, "%s: corrupt PNG image data", path);
        return false;
    }
    std::vector<uint8_t> px(size_t(h) * stride), zero(stride, 0);
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t ft = raw[size_t(y) * (stride + 1)];
        const uint8_t *in = &raw[size_t(y) * (stride + 1) + 1];
        uint8_t *cur = &px[size_t(y) * stride];
        const uint8_t *prev = y ? &px[size_t(y - 1) * stride] : zero.data();
        for (size_t x = 0; x < stride; ++x) {
            const int a = x >= size_t(channels) ? cur[x - channels] : 0;
            const int b = prev[x];
            const int c = x >= size_t(channels) ? prev[x - channels] : 0;
            int pred = 0;
            if (ft == 1) pred = a;
            else if (ft == 2) pred = b;
            else if (ft == 3) pred = (a + b) / 2;
            else if (ft == 4) {
                const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
                pred = (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
            } else if (ft != 0) {
                set_error(error, error_len, "%s: bad PNG row filter %u", path, unsigned(ft));
                return false;
            }
            cur[x] = uint8_t(in[x] + pred);
        }
    }
    img.width = w;
    img.height = h;
    img.rgba.assign(size_t(w) * h * 4, 255);
    for (size_t i = 0; i < size_t(w) * h; ++i) {
        const uint8_t *s = &px[i * channels];
        uint8_t *d = &img.rgba[i * 4];
        if (channels <= 2) {
            d[0] = d[1] = d[2] = s[0];
            if (channels == 2) d[3] = s[1];
        } else {
            d[0] = s[0]; d[1] = s[1]; d[2] = s[2];
            if (channels == 4) d[3] = s[3];
        }
    }
    return true;
}

// Luma (Rec. 709 weights on 8-bit values), metal_bridge.mm:2383-2386.
double luma_of(const Image &im, size_t p) {
    const uint8_t *c = &im.rgba[p * 4];
    return 0.2126 * double(c[0]) + 0.7152 * double(c[1]) + 0.0722 * double(c[2]);
}

// SSIM of two luma sets with sample (n - 1) statistics and the 8-bit
// constants C1 = (0.01 * 255)^2, C2 = (0.03 * 255)^2.
double ssim_of(const double *a, const double *b, size_t n, bool empty_is_one) {
    const double c1 = 6.5025, c2 = 58.5225;
    double ma = 0, mb = 0;
    for (size_t i = 0; i < n; ++i) {
        ma += a[i];
        mb += b[i];
    }
    ma /= double(n);
    mb /= double(n);
    double va = 0, vb = 0, cv = 0;
    for (size_t i = 0; i < n; ++i) {
        va += (a[i] - ma) * (a[i] - ma);
        vb += (b[i] - mb) * (b[i] - mb);
        cv += (a[i] - ma) * (b[i] - mb);
    }
    const double den = std::max(1.0, double(n) - 1.0);
    va /= den;
    vb /= den;
    cv /= den;
    const double num = (2.0 * ma * mb + c1) * (2.0 * cv + c2);
    const double dnm = (ma * ma + mb * mb + c1) * (va + vb + c2);
    return dnm > 0.0 ? num / dnm : (empty_is_one ? 1.0 : 0.0);
}

// Pixel-count-weighted mean SSIM over 8x8 luma blocks (metal_bridge.mm:
// 2388-2436).
double block_ssim(const Image &a, const Image &b) {
    double sum = 0, weight = 0;
    std::vector<double> la, lb;
    for (uint32_t by = 0; by < a.height; by += 8)
        for (uint32_t bx = 0; bx < a.width; bx += 8) {
            la.clear();
            lb.clear();
            for (uint32_t y = by; y < std::min(a.height, by + 8u); ++y)
                for (uint32_t x = bx; x < std::min(a.width, bx + 8u); ++x) {
                    la.push_back(luma_of(a, size_t(y) * a.width + x));
                    lb.push_back(luma_of(b, size_t(y) * a.width + x));
                }
            const double n = double(la.size());
            sum += ssim_of(la.data(), lb.data(), la.size(), true) * n;
            weight += n;
        }
    return weight > 0 ? std::clamp(sum / weight, -1.0, 1.0) : 0.0;
}

// SSIM of the 64x64 box-downsampled lumas (metal_bridge.mm:2438-2497).
double low_frequency_ssim(const Image &a, const Image &b) {
    const uint32_t T = 64;
    std::vector<double> da(T * T), db(T * T);
    for (uint32_t y = 0; y < T; ++y) {
        const uint32_t y0 = uint32_t(uint64_t(y) * a.height / T);
        const uint32_t y1 = std::max(uint32_t(uint64_t(y + 1) * a.height / T), y0 + 1u);
        for (uint32_t x = 0; x < T; ++x) {
            const uint32_t x0 = uint32_t(uint64_t(x) * a.width / T);
            const uint32_t x1 = std::max(uint32_t(uint64_t(x + 1) * a.width / T), x0 + 1u);
            double sa = 0, sb = 0, n = 0;
            for (uint32_t sy = y0; sy < y1; ++sy)
                for (uint32_t sx = x0; sx < x1; ++sx) {
                    sa += luma_of(a, size_t(sy) * a.width + sx);
                    sb += luma_of(b, size_t(sy) * a.width + sx);
                    n += 1;
                }
            da[y * T + x] = sa / n;
            db[y * T + x] = sb / n;
        }
    }
    return std::clamp(ssim_of(da.data(), db.data(), da.size(), true), -1.0, 1.0);
}

struct CompareGate {
    const char *name;
    double mae, strict_mae, strict_ssim;
};

// Scene-class gates chosen from the candidate path (metal_bridge.mm:
// 2358-2381).
CompareGate gate_for(const char *path) {
    const std::string p(path ? path : "");
    auto has = [&](const char *k) { return p.find(k) != std::string::npos; };
    const bool preview = has("preview-parity");
    if (preview && has("Ball_Fractal")) return {"preview_ball_fractal", 50.0, 35.0, 0.50};
    if (preview && has("Glass_Ball")) return {"preview_glass", 60.0, 42.0, 0.45};
    if (preview) return {"preview", 90.0, 55.0, 0.45};
    static const CompareGate named[] = {
        {"cornell", 50.0, 30.0, 0.55},      {"glass", 50.0, 30.0, 0.30},          {"ball_fractal", 45.0, 28.0, 0.55},
        {"cage_fractal", 70.0, 45.0, 0.45}, {"ifs_fractal", 75.0, 48.0, 0.45},    {"mandelbox_fractal", 80.0, 52.0, 0.42},
        {"menger_sponge", 60.0, 40.0, 0.48}, {"tower_fractal", 75.0, 48.0, 0.45}, {"tree_fractal", 60.0, 38.0, 0.48}};
    static const char *keys[] = {"Cornell_Box", "Glass_Ball", "Ball_Fractal", "Cage_Fractal", "IFS_Fractal",
                                 "Mandelbox_Fractal", "Menger_Sponge", "Tower_Fractal", "Tree_Fractal"};
    for (int i = 0; i < 9; ++i)
        if (has(keys[i])) return named[i];
    return {"generic", 90.0, 55.0, 0.45};
}

// Three panels (baseline | candidate | difference heat map), 360 px wide
// each, 12 px gaps on a grey 32 background (metal_bridge.mm:2315-2356).
std::vector<uint8_t> comparison_sheet(const Image &a, const Image &b, uint32_t &ow, uint32_t &oh) {
    const uint32_t gap = 12, pw = 360;
    const uint32_t ph = std::max<uint32_t>(1u, uint32_t(std::lround(double(a.height) * pw / double(a.width))));
    ow = pw * 3 + gap * 2;
    oh = ph;
    std::vector<uint8_t> sheet(size_t(ow) * oh * 4, 32);
    for (uint32_t y = 0; y < oh; ++y) {
        const uint32_t sy = std::min<uint32_t>(a.height - 1, uint32_t(uint64_t(y) * a.height / ph));
        for (uint32_t x = 0; x < pw; ++x) {
            const uint32_t sx = std::min<uint32_t>(a.width - 1, uint32_t(uint64_t(x) * a.width / pw));
            const size_t src = (size_t(sy) * a.width + sx) * 4;
            uint8_t *pa = &sheet[(size_t(y) * ow + x) * 4];
            uint8_t *pb = &sheet[(size_t(y) * ow + pw + gap + x) * 4];
            uint8_t *pd = &sheet[(size_t(y) * ow + 2 * (pw + gap) + x) * 4];
            std::memcpy(pa, &a.rgba[src], 4);
            std::memcpy(pb, &b.rgba[src], 4);
            int d = 0;
            for (int ch = 0; ch < 3; ++ch) d += std::abs(int(a.rgba[src + ch]) - int(b.rgba[src + ch]));
            const uint8_t heat = uint8_t(std::min(255, d / 3 * 4));
            pd[0] = heat;
            pd[1] = heat / 2;
            pd[2] = uint8_t(255 - heat);
            pd[3] = 255;
        }
    }
    return sheet;
}

// metal_bridge.mm:2302-2313: distinct colours among the first 2048 found.
size_t unique_colour_sample(const std::vector<uint8_t> &rgba) {
    std::vector<uint32_t> colours;
    for (size_t i = 0; i + 3 < rgba.size(); i += 4) {
        uint32_t c;
        std::memcpy(&c, &rgba[i], 4);
        if (std::find(colours.begin(), colours.end(), c) == colours.end()) {
            colours.push_back(c);
            if (colours.size() >= 2048) break;
        }
    }
    return colours.size();
}

}  // namespace

// ---------------------------------------------------------------------------
// Render plan.
// ---------------------------------------------------------------------------
struct FptHipPlan {
    FptRenderConfig host;
    FptRenderConfig *dcfg = nullptr;
    fpt::VoxelStore voxels;  // populated for the voxel backend
    double build_ms = 0.0;
    // SDF launch shape (fpt_hip_plan_tune).
    uint32_t kernel = 1;         // 0 one path per lane, 1 persistent
    uint32_t workgroups = 0;     // persistent grid; 0 = 2 per compute unit
    uint32_t event_batch = 0;    // waiting paths before per-hit work runs (0: the instance's default)
    // Path-stash refill: non-marching lanes of a wave before they exchange
    // their paths for parked marching ones (fpt_hip_plan_tune_stash).
    uint32_t refill = fpt::kDefaultRefill;
    // Sample runs per pixel of the single-scene persistent kernels
    // (fpt_hip_plan_tune_stash); their per-sample colours and run counters.
    uint32_t sample_groups = 8;
    float *split_colors = nullptr;
    uint32_t *split_done = nullptr;
    size_t split_colors_len = 0;
    uint32_t *counter = nullptr; // persistent pixel counter
    // Hand-out order of the persistent kernel's 8x8 tiles: the most work
    // first (longest-processing-time order from a probe of primary rays), so
    // the frame's tail is made of cheap tiles and lanes finish together.
    uint32_t *tile_order = nullptr;
    uint32_t tile_order_mode = 1;  // 1 longest-first, 0 row-major (fpt_hip_set_tile_order)
    fpt::ProposerHost proposer;  // learned stepping (fpt_hip_plan_set_proposer)
    // Topology-specialised persistent kernel (fpt_hip_plan_specialize).
    std::shared_ptr<fpt::SpecializedModule> spec;
    fpt::SpecializeReport spec_report;
};

namespace {
uint32_t default_workgroups() {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    return uint32_t(cus > 0 ? cus : 256) * 2u;
}

// Shortest sample run: every finished run publishes its colours with a
// device-scope release (an L2 write-back on gfx950) and the last one
// acquires them (an L1/L2 invalidate), so runs must be long enough for that
// to be rare.  With write-through colour stores, runs of 8 sped the 16-spp
// typed program 8% and slowed the 28-primitive union 4% (the autotuner
// chooses for such frames); runs of 4 were slower still.
constexpr uint32_t kMinSampleRun = 8u;

// Sample runs of a launch over samples [start, start + chunk) of the frame:
// runs of ceil(n / groups) samples, every run non-empty and at least
// kMinSampleRun long.  Single-scene exact-march plans only (no proposer;
// imported Mandelbulber scenes included), frames whose sides are multiples of 8, and
// only while the per-sample colours fit in 8 GB; otherwise one slot per
// pixel.  Allocates the colour and counter
// buffers on first use.
hipError_t plan_split(FptHipPlan *plan, uint32_t start, uint32_t chunk, fpt::SampleSplit &split) {
    split = fpt::SampleSplit{nullptr, nullptr, 1u, 0u, 0u};
    const FptRenderConfig &c = plan->host;
    const uint32_t spp = clamp_samples(c.samples);
    const uint32_t n = std::min(start + chunk, spp) > start ? std::min(start + chunk, spp) - start : 0u;
    if (plan->sample_groups <= 1u || n <= 1u || plan->proposer.enabled)
        return hipSuccess;
    // Frames whose edge tiles carry padding render one slot per pixel: sample
    // runs are verified only on frames that tile exactly into 8x8.
    if (c.width % 8u != 0u || c.height % 8u != 0u) return hipSuccess;
    const size_t npix = size_t(c.width) * c.height, len = npix * spp;
    if (len * 3u * sizeof(float) > (size_t(8) << 30)) return hipSuccess;
    const uint32_t groups = std::min(plan->sample_groups, n / kMinSampleRun);
    if (groups <= 1u) return hipSuccess;
    const uint32_t run = (n + groups - 1u) / groups;
    if (plan->split_colors_len < len) {
        if (plan->split_colors) (void)hipFree(plan->split_colors);
        if (plan->split_done) (void)hipFree(plan->split_done);
        plan->split_colors = nullptr;
        plan->split_done = nullptr;
        plan->split_colors_len = 0;
        hipError_t e = hipMalloc(&plan->split_colors, len * 3u * sizeof(float));
        if (e == hipSuccess) e = hipMalloc(&plan->split_done, npix * sizeof(uint32_t));
        if (e != hipSuccess) return e;
        plan->split_colors_len = len;
    }
    split.colors = plan->split_colors;
    split.done = plan->split_done;
    split.groups = (n + run - 1u) / run;
    split.chunk = run;
    split.stride = spp;
    return hipSuccess;
}

hipError_t launch_sdf(FptHipPlan *plan, float4 *acc, uint32_t start, uint32_t chunk, fpt::DeviceProfile *prof,
                      hipStream_t s) {
    const FptRenderConfig &c = plan->host;
    if (plan->kernel == 0u) return fpt::launch_sdf_accumulate(plan->dcfg, c, acc, start, chunk, prof, s);
    if (!plan->counter) {
        hipError_t e = hipMalloc(&plan->counter, sizeof(uint32_t));
        if (e != hipSuccess) return e;
    }
    fpt::SampleSplit split;
    hipError_t e = plan_split(plan, start, chunk, split);
    if (e != hipSuccess) return e;
    // Imported Mandelbulber scenes run the stash-free instance: as many
    // workgroups per compute unit as its waves per SIMD.
    const uint32_t wg = plan->workgroups ? plan->workgroups
                        : c.sdf_id == FPT_SDF_MANDELBULBER
                            ? default_workgroups() / 2u * uint32_t(fpt::kMandelbulberPersistentWaves)
                            : default_workgroups();
    if (plan->spec && !plan->proposer.enabled && !prof && !fpt::single_bulb_program(c)) {
        // Same kernel, built for the program's topology (counting passes use
        // the interpreter: the work counts are the same).
        e = fpt::launch_zero_counter(plan->counter, s);
        if (e == hipSuccess && split.groups > 1u) e = fpt::launch_zero_words(split.done, c.width * c.height, s);
        if (e != hipSuccess) return e;
        return fpt::launch_specialized(*plan->spec, plan->dcfg, acc, start, chunk, plan->counter,
                                       fpt::persistent_slots(c.width, c.height) * split.groups, wg,
                                       plan->event_batch ? int(plan->event_batch) : fpt::default_event_batch(true, split.groups > 1u),
                                       int(plan->refill), plan->tile_order_mode ? plan->tile_order : nullptr, split, s);
    }
    return fpt::launch_sdf_persistent(plan->dcfg, c, acc, start, chunk, plan->counter, wg, int(plan->event_batch),
                                      int(plan->refill), prof, plan->proposer.enabled ? &plan->proposer.dev : nullptr,
                                      plan->tile_order_mode ? plan->tile_order : nullptr, split, s);
}

// Hand-out order of tiles from their work estimates (scene_tile_costs):
// most work first, ties in row-major order.  Cheap tiles then form the
// frame's tail, so lanes run out of work together.  Any order gives the same
// image, since every pixel is rendered by one lane alone.
std::vector<uint32_t> tile_order_from_costs(const std::vector<float> &cost) {
    std::vector<uint32_t> order(cost.size());
    for (uint32_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
    return order;
}

// The persistent kernel's tile order (tile_order_from_costs).
// Work estimate of each 8x8 tile of one scene (tiles row-major) into
// cost[tile]: one whole path sample per pixel (launch_pixel_cost), summed
// over the tile's pixels.
hipError_t scene_tile_costs(const FptRenderConfig *dcfg, const FptRenderConfig &c, float *cost) {
    const uint32_t tiles_x = (c.width + 7u) / 8u;
    const size_t npix = size_t(c.width) * c.height;
    DevBuf px_dev;
    hipError_t e = px_dev.alloc(npix * sizeof(float));
    if (e == hipSuccess) e = fpt::launch_pixel_cost(dcfg, c, px_dev.as<float>(), nullptr);
    std::vector<float> px(npix);
    if (e == hipSuccess) e = hipMemcpy(px.data(), px_dev.p, npix * sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return e;
    for (uint32_t y = 0; y < c.height; ++y)
        for (uint32_t x = 0; x < c.width; ++x) cost[(y / 8u) * tiles_x + x / 8u] += px[size_t(y) * c.width + x];
    return hipSuccess;
}

int plan_tile_order(FptHipPlan *plan, char *error, size_t error_len) {
    const FptRenderConfig &c = plan->host;
    const uint32_t tiles_x = (c.width + 7u) / 8u, ntiles = tiles_x * ((c.height + 7u) / 8u);
    std::vector<float> h(ntiles, 0.0f);
    HIP_TRY(scene_tile_costs(plan->dcfg, c, h.data()), "tile cost probe");
    const std::vector<uint32_t> order = tile_order_from_costs(h);
    HIP_TRY(hipMalloc(&plan->tile_order, size_t(ntiles) * sizeof(uint32_t)), "tile order allocation");
    HIP_TRY(hipMemcpy(plan->tile_order, order.data(), size_t(ntiles) * sizeof(uint32_t), hipMemcpyHostToDevice),
            "tile order upload");
    return 0;
}

// Attach a validated topology-specialised kernel to `plan`
// (2^20 validation points; fails without attaching when any point fails).
int plan_specialize(FptHipPlan *plan, char *error, size_t error_len) {
    std::shared_ptr<fpt::SpecializedModule> mod;
    fpt::SpecializeReport rep;
    std::string failure;
    if (fpt::specialize(plan->host, mod, rep, failure)) {
        plan->spec_report = rep;
        set_error(error, error_len, "%s", failure.c_str());
        return 1;
    }
    Stream st;
    HIP_TRY(hipStreamCreate(&st.s), "hipStreamCreate");
    HIP_TRY(fpt::validate_specialized(*mod, plan->dcfg, 1u << 20, rep, st.s), "topology validation");
    plan->spec_report = rep;
    if (rep.distance_failures || rep.gradient_failures) {
        set_error(error, error_len,
                  "topology-specialised program failed validation: %llu distance and %llu gradient failures of %llu "
                  "(max errors %g, %g)",
                  (unsigned long long)rep.distance_failures, (unsigned long long)rep.gradient_failures,
                  (unsigned long long)rep.sample_count, double(rep.max_distance_error),
                  double(rep.max_gradient_error));
        return 1;
    }
    plan->spec = mod;
    return 0;
}

void fill_specialization_stats(const FptHipPlan *plan, FptSpecializationStats *st) {
    const fpt::SpecializeReport &r = plan->spec_report;
    std::memset(st, 0, sizeof(*st));
    st->cache_status = r.cache_status;
    st->instruction_count = std::min<uint32_t>(plan->host.sdf_program_count, FPT_SDF_PROGRAM_MAX_OPS);
    st->source_bytes = r.source_bytes;
    st->code_bytes = r.code_bytes;
    st->sample_count = r.sample_count;
    st->distance_failures = r.distance_failures;
    st->gradient_failures = r.gradient_failures;
    st->max_distance_error = r.max_distance_error;
    st->max_gradient_error = r.max_gradient_error;
    st->compile_ms = r.compile_ms;
    st->link_ms = r.link_ms;
    st->validate_ms = r.validate_ms;
}
// Samples per SDF launch.  The one-path-per-lane kernel follows the
// reference's accumulation policy (chunk_for); the persistent kernel keeps
// every pixel's running mean in a register across its samples, so one launch
// covers the whole frame with the same sequence of mix() steps (the float4
// round trip between chunked launches is exact): the image is bit-identical
// and the accumulation buffer is read and written once per frame.
uint32_t sdf_launch_chunk(const FptHipPlan *plan) {
    const uint32_t spp = clamp_samples(plan->host.samples);
    return plan->kernel == 1u ? spp : chunk_for(plan->host);
}
}  // namespace

extern "C" {

const char *fpt_hip_last_error(void) { return g_last_error.c_str(); }

int fpt_hip_plan_create(const FptRenderConfig *config, FptHipPlan **plan, char *error, size_t error_len) {
    if (!plan) {
        set_error(error, error_len, "missing plan out-pointer");
        return 1;
    }
    *plan = nullptr;
    if (validate_config(config, error, error_len)) return 1;
    if (check_device(error, error_len)) return 1;
    // The HDRI LUT always comes from the scene's HDRI file (metal_bridge.mm:
    // 4330-4332): the plan works on the hydrated copy.
    FptRenderConfig hydrated = *config;
    if (!fpt::hydrate_hdri(hydrated, error, error_len)) return 1;
    config = &hydrated;
    FptHipPlan *p = new FptHipPlan();
    p->host = *config;
    hipError_t e = hipMalloc(&p->dcfg, sizeof(FptRenderConfig));
    if (e == hipSuccess) e = hipMemcpy(p->dcfg, config, sizeof(FptRenderConfig), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        set_error(error, error_len, "config upload failed: %s", hipGetErrorString(e));
        fpt_hip_plan_destroy(p);
        return 1;
    }
    if (config->renderer_backend == FPT_RENDERER_SDF && config->camera_dof > 0.0f && !(config->focus_distance > 0.0f)) {
        // Focus prepass (metal_bridge.mm:4827-4838, 5455-5481): the estimate
        // replaces the unset focus distance in the device config and in the
        // plan's host copy.
        e = fpt::launch_focus_prepass(p->dcfg, 1u, nullptr);
        if (e == hipSuccess)
            e = hipMemcpy(&p->host.focus_distance, &p->dcfg->focus_distance, sizeof(float), hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            set_error(error, error_len, "focus-distance prepass failed: %s", hipGetErrorString(e));
            fpt_hip_plan_destroy(p);
            return 1;
        }
    }
    if (config->renderer_backend == FPT_RENDERER_SDF && plan_tile_order(p, error, error_len)) {
        fpt_hip_plan_destroy(p);
        return 1;
    }
    if (config->renderer_backend == FPT_RENDERER_VOXEL) {
        auto t0 = std::chrono::steady_clock::now();
        if (fpt::voxel_build(*config, p->dcfg, p->voxels, nullptr, error, error_len)) {
            g_last_error = error && error_len ? error : "voxel build failed";
            fpt_hip_plan_destroy(p);
            return 1;
        }
        p->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    *plan = p;
    return 0;
}

int fpt_hip_hydrate_hdri(FptRenderConfig *config, char *error, size_t error_len) {
    if (!config) {
        set_error(error, error_len, "missing render config");
        return 1;
    }
    return fpt::hydrate_hdri(*config, error, error_len) ? 0 : 1;
}

int fpt_hip_plan_render(FptHipPlan *plan, float *device_accum, void *stream, char *error, size_t error_len) {
    if (!plan || !device_accum) {
        set_error(error, error_len, "missing plan or accumulation buffer");
        return 1;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const FptRenderConfig &c = plan->host;
    const uint32_t spp = clamp_samples(c.samples);
    const uint32_t chunk = c.renderer_backend == FPT_RENDERER_VOXEL ? chunk_for(c) : sdf_launch_chunk(plan);
    float4 *acc = reinterpret_cast<float4 *>(device_accum);
    for (uint32_t start = 0; start < spp; start += chunk) {
        hipError_t e = c.renderer_backend == FPT_RENDERER_VOXEL
                           ? fpt::launch_voxel_accumulate(plan->dcfg, c, plan->voxels, acc, start, chunk, s, nullptr,
                                                          plan->kernel == 1u, plan->workgroups, int(plan->event_batch))
                           : launch_sdf(plan, acc, start, chunk, nullptr, s);
        HIP_TRY(e, "render kernel launch");
    }
    return 0;
}

int fpt_hip_plan_tune(FptHipPlan *plan, uint32_t kernel, uint32_t workgroups, uint32_t event_batch, char *error,
                      size_t error_len) {
    if (!plan) {
        set_error(error, error_len, "missing plan");
        return 1;
    }
    if (kernel > 1u || event_batch > 128u || workgroups > 65536u) {
        set_error(error, error_len, "invalid launch shape (kernel %u, workgroups %u, event_batch %u)", kernel,
                  workgroups, event_batch);
        return 1;
    }
    if (kernel == 0u && plan->proposer.enabled) {
        set_error(error, error_len, "learned stepping runs on the persistent kernel (kernel 1)");
        return 1;
    }
    plan->kernel = kernel;
    plan->workgroups = workgroups;
    plan->event_batch = event_batch;
    return 0;
}

int fpt_hip_plan_tune_stash(FptHipPlan *plan, uint32_t refill, uint32_t sample_groups, char *error,
                            size_t error_len) {
    if (!plan) {
        set_error(error, error_len, "missing plan");
        return 1;
    }
    if (refill < 1u || refill > 64u) {
        set_error(error, error_len, "invalid stash refill threshold %u (1..64 lanes)", refill);
        return 1;
    }
    if (sample_groups < 1u || sample_groups > 64u) {
        set_error(error, error_len, "invalid sample run count %u (1..64 per pixel)", sample_groups);
        return 1;
    }
    plan->refill = refill;
    plan->sample_groups = sample_groups;
    return 0;
}

void fpt_hip_plan_destroy(FptHipPlan *plan) {
    if (!plan) return;
    if (plan->dcfg) (void)hipFree(plan->dcfg);
    if (plan->counter) (void)hipFree(plan->counter);
    if (plan->split_colors) (void)hipFree(plan->split_colors);
    if (plan->split_done) (void)hipFree(plan->split_done);
    if (plan->tile_order) (void)hipFree(plan->tile_order);
    fpt::voxel_free(plan->voxels);
    fpt::proposer_release(plan->proposer);
    plan->spec.reset();
    delete plan;
}

int fpt_hip_plan_set_proposer(FptHipPlan *plan, const float *params, size_t param_count,
                              const FptProposerSettings *settings, char *error, size_t error_len) {
    if (!plan) {
        set_error(error, error_len, "missing plan");
        return 1;
    }
    if (!settings) {
        fpt::proposer_release(plan->proposer);
        return 0;
    }
    if (plan->host.renderer_backend != FPT_RENDERER_SDF) {
        set_error(error, error_len, "learned stepping applies to the SDF backend");
        return 1;
    }
    if (reject_mandelbulber(&plan->host, "learned stepping", error, error_len)) return 1;
    if (plan->kernel == 0u) {
        set_error(error, error_len, "learned stepping runs on the persistent kernel (kernel 1)");
        return 1;
    }
    if (const char *msg = fpt::proposer_upload(plan->proposer, params, param_count, settings->box_half,
                                               settings->switch_distance, settings->resume_distance,
                                               settings->step_scale)) {
        set_error(error, error_len, "%s", msg);
        return 1;
    }
    return 0;
}

int fpt_hip_plan_voxel_stats(FptHipPlan *plan, float *device_accum, FptVoxelTraceStats *stats, char *error,
                             size_t error_len) {
    if (!plan || !device_accum || !stats) {
        set_error(error, error_len, "missing plan, accumulation buffer or stats");
        return 1;
    }
    if (plan->host.renderer_backend != FPT_RENDERER_VOXEL) {
        set_error(error, error_len, "the voxel trace counters apply to the voxel backend");
        return 1;
    }
    std::memset(stats, 0, sizeof(*stats));
    Stream st;
    HIP_TRY(hipStreamCreate(&st.s), "hipStreamCreate");
    DevBuf counts;
    HIP_TRY(counts.alloc(3 * sizeof(unsigned long long)), "counter buffer");
    HIP_TRY(hipMemsetAsync(counts.p, 0, 3 * sizeof(unsigned long long), st.s), "counter clear");
    const FptRenderConfig &c = plan->host;
    const uint32_t spp = clamp_samples(c.samples), chunk = chunk_for(c);
    float4 *acc = reinterpret_cast<float4 *>(device_accum);
    for (uint32_t s0 = 0; s0 < spp; s0 += chunk)
        HIP_TRY(fpt::launch_voxel_accumulate(plan->dcfg, c, plan->voxels, acc, s0, chunk, st.s,
                                             counts.as<unsigned long long>(), plan->kernel == 1u, plan->workgroups,
                                             int(plan->event_batch)),
                "voxel counting launch");
    unsigned long long h[3] = {0, 0, 0};
    HIP_TRY(hipMemcpyAsync(h, counts.p, sizeof(h), hipMemcpyDeviceToHost, st.s), "counter readback");
    HIP_TRY(hipStreamSynchronize(st.s), "voxel counting pass");
    stats->visited_cells = h[0];
    stats->page_reads = h[1];
    stats->cell_reads = h[2];
    stats->pixels = uint64_t(c.width) * c.height;
    return 0;
}

}  // extern "C"

namespace {
// The untimed counting pass of a plan's full frame (the PROFILE kernels).
int profile_pass(FptHipPlan *plan, float *device_accum, fpt::DeviceProfile &hp, char *error, size_t error_len) {
    if (plan->host.renderer_backend != FPT_RENDERER_SDF) {
        set_error(error, error_len, "the counting pass applies to the SDF backend");
        return 1;
    }
    Stream st;
    HIP_TRY(hipStreamCreate(&st.s), "hipStreamCreate");
    DevBuf prof;
    HIP_TRY(prof.alloc(sizeof(fpt::DeviceProfile)), "profile buffer");
    HIP_TRY(hipMemsetAsync(prof.p, 0, sizeof(fpt::DeviceProfile), st.s), "profile clear");
    const uint32_t spp = clamp_samples(plan->host.samples), chunk = sdf_launch_chunk(plan);
    float4 *acc = reinterpret_cast<float4 *>(device_accum);
    for (uint32_t s0 = 0; s0 < spp; s0 += chunk)
        HIP_TRY(launch_sdf(plan, acc, s0, chunk, prof.as<fpt::DeviceProfile>(), st.s), "profile kernel launch");
    HIP_TRY(hipMemcpyAsync(&hp, prof.p, sizeof(hp), hipMemcpyDeviceToHost, st.s), "profile readback");
    HIP_TRY(hipStreamSynchronize(st.s), "profile pass");
    return 0;
}
}  // namespace

extern "C" {

int fpt_hip_plan_profile(FptHipPlan *plan, float *device_accum, FptSdfProfileStats *stats,
                         FptProposerStats *learned, char *error, size_t error_len) {
    if (!plan || !device_accum || !stats) {
        set_error(error, error_len, "missing plan, accumulation buffer or stats");
        return 1;
    }
    std::memset(stats, 0, sizeof(*stats));
    if (learned) std::memset(learned, 0, sizeof(*learned));
    fpt::DeviceProfile hp;
    if (profile_pass(plan, device_accum, hp, error, error_len)) return 1;
    stats->primary_steps = hp.primary_steps;
    stats->secondary_steps = hp.secondary_steps;
    stats->shadow_steps = hp.shadow_steps;
    stats->normal_evals = hp.normal_evals;
    stats->bounces = hp.bounces;
    stats->pixels = hp.pixels;
    stats->distance_evals = hp.distance_evals;
    stats->march_orbit_iterations = hp.orbit_iterations;
    if (learned) {
        learned->proposer_calls = hp.proposer_calls;
        learned->proposals = hp.proposals;
        learned->far_steps = hp.far_steps;
        learned->backtracks = hp.backtracks;
    }
    return 0;
}

int fpt_hip_plan_lane_stats(FptHipPlan *plan, float *device_accum, FptLaneStats *stats, char *error,
                            size_t error_len) {
    if (!plan || !device_accum || !stats) {
        set_error(error, error_len, "missing plan, accumulation buffer or stats");
        return 1;
    }
    if (plan->kernel != 1u) {
        set_error(error, error_len, "lane statistics are kept by the persistent kernel (kernel 1)");
        return 1;
    }
    std::memset(stats, 0, sizeof(*stats));
    fpt::DeviceProfile hp;
    if (profile_pass(plan, device_accum, hp, error, error_len)) return 1;
    static_assert(sizeof(FptLaneStats) == sizeof(hp.lane), "lane counters");
    std::memcpy(stats, hp.lane, sizeof(hp.lane));
    return 0;
}

int fpt_hip_plan_specialize(FptHipPlan *plan, uint32_t enable, FptSpecializationStats *stats, char *error,
                            size_t error_len) {
    if (!plan) {
        set_error(error, error_len, "missing plan");
        return 1;
    }
    if (!enable) {
        plan->spec.reset();
        plan->spec_report = fpt::SpecializeReport();
        if (stats) std::memset(stats, 0, sizeof(*stats));
        return 0;
    }
    if (plan->host.renderer_backend != FPT_RENDERER_SDF) {
        set_error(error, error_len, "runtime-source compilation requires the direct SDF renderer");
        return 1;
    }
    const int rc = plan_specialize(plan, error, error_len);
    if (stats) fill_specialization_stats(plan, stats);
    return rc;
}

int fpt_hip_plan_specialized_surface(FptHipPlan *plan, const float *points_xyzw, size_t point_count,
                                     float *out_dgrad, char *error, size_t error_len) {
    if (!plan || !points_xyzw || !out_dgrad) {
        set_error(error, error_len, "missing plan, points or output buffer");
        return 1;
    }
    if (!plan->spec) {
        set_error(error, error_len, "the plan has no topology-specialised module (fpt_hip_plan_specialize)");
        return 1;
    }
    if (point_count == 0) return 0;
    if (point_count > 0xffffffffull) {
        set_error(error, error_len, "too many points (%zu)", point_count);
        return 1;
    }
    Stream st;
    HIP_TRY(hipStreamCreate(&st.s), "hipStreamCreate");
    DevBuf dp, dout;
    HIP_TRY(dp.alloc(point_count * 16), "point buffer allocation");
    HIP_TRY(dout.alloc(point_count * 16), "output buffer allocation");
    HIP_TRY(hipMemcpyAsync(dp.p, points_xyzw, point_count * 16, hipMemcpyHostToDevice, st.s), "point upload");
    HIP_TRY(fpt::sample_specialized(*plan->spec, plan->dcfg, dp.as<float4>(), uint32_t(point_count),
                                    dout.as<float4>(), st.s),
            "specialised surface kernel launch");
    HIP_TRY(hipMemcpyAsync(out_dgrad, dout.p, point_count * 16, hipMemcpyDeviceToHost, st.s), "readback");
    HIP_TRY(hipStreamSynchronize(st.s), "specialised surface sampling");
    return 0;
}

int fpt_hip_topology_source(const FptRenderConfig *config, char *out, size_t out_len, size_t *needed, char *error,
                            size_t error_len) {
    if (!config) {
        set_error(error, error_len, "missing render config");
        return 1;
    }
    std::string failure;
    const std::string src = fpt::topology_source(*config, failure);
    if (src.empty()) {
        set_error(error, error_len, "%s", failure.c_str());
        return 1;
    }
    if (needed) *needed = src.size() + 1;
    if (out && out_len) {
        if (out_len < src.size() + 1) {
            set_error(error, error_len, "source buffer is too small (%zu < %zu)", out_len, src.size() + 1);
            return 1;
        }
        std::memcpy(out, src.c_str(), src.size() + 1);
    }
    return 0;
}

int fpt_hip_proposer_samples(const FptRenderConfig *config, float box_half, float band, uint32_t count,
                             uint64_t seed, uint32_t max_attempts, float *out_xyzd, uint32_t *failures,
                             char *error, size_t error_len) {
    if (validate_config(config, error, error_len)) return 1;
    if (reject_mandelbulber(config, "learned stepping", error, error_len)) return 1;
    if (!out_xyzd || !failures) {
        set_error(error, error_len, "missing sample or failure-count buffer");
        return 1;
    }
    if (!(box_half > 0.0f) || !(band > 0.0f) || max_attempts == 0u) {
        set_error(error, error_len, "box_half and band must be positive and max_attempts non-zero");
        return 1;
    }
    *failures = 0;
    if (count == 0) return 0;
    if (check_device(error, error_len)) return 1;
    Stream st;
    HIP_TRY(hipStreamCreate(&st.s), "hipStreamCreate");
    DevBuf dcfg, dout, dfail;
    HIP_TRY(dcfg.alloc(sizeof(FptRenderConfig)), "config allocation");
    HIP_TRY(dout.alloc(size_t(count) * sizeof(float4)), "sample buffer allocation");
    HIP_TRY(dfail.alloc(sizeof(uint32_t)), "counter allocation");
    HIP_TRY(hipMemcpyAsync(dcfg.p, config, sizeof(FptRenderConfig), hipMemcpyHostToDevice, st.s), "config upload");
    HIP_TRY(hipMemsetAsync(dfail.p, 0, sizeof(uint32_t), st.s), "counter clear");
    HIP_TRY(fpt::launch_proposer_samples(dcfg.as<FptRenderConfig>(), box_half, band, count, seed, max_attempts,
                                         dout.as<float4>(), dfail.as<uint32_t>(), st.s),
            "sample kernel launch");
    HIP_TRY(hipMemcpyAsync(out_xyzd, dout.p, size_t(count) * sizeof(float4), hipMemcpyDeviceToHost, st.s),
            "sample readback");
    HIP_TRY(hipMemcpyAsync(failures, dfail.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st.s), "counter readback");
    HIP_TRY(hipStreamSynchronize(st.s), "sampling");
    return 0;
}

int fpt_hip_proposer_eval(const float *params, size_t param_count, float box_half, const float *points_xyzw,
                          size_t point_count, float *out, char *error, size_t error_len) {
    if (!points_xyzw || !out) {
        set_error(error, error_len, "missing point or output buffer");
        return 1;
    }
    if (check_device(error, error_len)) return 1;
    fpt::ProposerHost ph;
    struct Release {
        fpt::ProposerHost &h;
        ~Release() { fpt::proposer_release(h); }
    } release{ph};
    if (const char *msg = fpt::proposer_upload(ph, params, param_count, box_half, 0.02f, 0.02f, 0.5f)) {
        set_error(error, error_len, "%s", msg);
        return 1;
    }
    if (point_count == 0) return 0;
    Stream st;
    HIP_TRY(hipStreamCreate(&st.s), "hipStreamCreate");
    DevBuf dp, dout;
    HIP_TRY(dp.alloc(point_count * sizeof(float4)), "point buffer allocation");
    HIP_TRY(dout.alloc(point_count * sizeof(float)), "output allocation");
    HIP_TRY(hipMemcpyAsync(dp.p, points_xyzw, point_count * sizeof(float4), hipMemcpyHostToDevice, st.s),
            "point upload");
    HIP_TRY(fpt::launch_proposer_eval(ph.dev, dp.as<float4>(), point_count, dout.as<float>(), st.s),
            "proposer kernel launch");
    HIP_TRY(hipMemcpyAsync(out, dout.p, point_count * sizeof(float), hipMemcpyDeviceToHost, st.s), "output readback");
    HIP_TRY(hipStreamSynchronize(st.s), "proposer evaluation");
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched multi-scene rendering.
// ---------------------------------------------------------------------------
struct FptHipBatch {
    std::vector<FptRenderConfig> host;
    std::vector<fpt::BatchScene> scenes;
    FptRenderConfig *dcfgs = nullptr;
    fpt::BatchScene *dscenes = nullptr;
    uint32_t *counter = nullptr;
    uint32_t *tile_order = nullptr;  // longest-first order of the batch's tiles
    uint32_t tile_order_mode = 1;    // 1 longest-first, 0 row-major
    uint32_t total_slots = 0;
    size_t pixels = 0;
    bool lowered = false;  // some scene carries lowered program tables
    uint32_t workgroups = 0;
    uint32_t event_batch = 0;  // 0: the instance's default
    uint32_t use_graph = 1;
    // Captured launch sequence (counter reset + persistent kernel) for one
    // output buffer.
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    float *graph_accum = nullptr;
    uint32_t graph_wg = 0, graph_eb = 0, graph_order = 0;
};

namespace {
void batch_drop_graph(FptHipBatch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_accum = nullptr;
}
}  // namespace

extern "C" {

int fpt_hip_batch_create(const FptRenderConfig *configs, uint32_t count, FptHipBatch **out, char *error,
                         size_t error_len) {
    if (!out) {
        set_error(error, error_len, "missing batch out-pointer");
        return 1;
    }
    *out = nullptr;
    if (!configs || count == 0 || count > 65536) {
        set_error(error, error_len, "invalid scene count %u (1..65536)", count);
        return 1;
    }
    for (uint32_t i = 0; i < count; ++i) {
        if (validate_config(&configs[i], error, error_len)) return 1;
        if (reject_mandelbulber(&configs[i], "batched rendering", error, error_len)) return 1;
        if (configs[i].renderer_backend != FPT_RENDERER_SDF) {
            set_error(error, error_len, "scene %u: batched rendering supports the SDF backend only", i);
            return 1;
        }
    }
    if (check_device(error, error_len)) return 1;
    FptHipBatch *b = new FptHipBatch();
    b->host.assign(configs, configs + count);
    for (uint32_t i = 0; i < count; ++i) {
        if (!fpt::hydrate_hdri(b->host[i], error, error_len)) {
            fpt_hip_batch_destroy(b);
            return 1;
        }
    }
    uint64_t slot = 0, off = 0;
    for (uint32_t i = 0; i < count; ++i) {
        fpt::BatchScene sc;
        sc.slot_begin = uint32_t(slot);
        sc.slots = ((configs[i].width + 7u) / 8u) * ((configs[i].height + 7u) / 8u) * 64u;
        sc.width = configs[i].width;
        sc.height = configs[i].height;
        sc.accum_offset = off;
        slot += sc.slots;
        off += uint64_t(configs[i].width) * configs[i].height;
        b->scenes.push_back(sc);
    }
    if (slot > 0xffffffffull) {
        set_error(error, error_len, "batch has too many pixels (%llu slots)", (unsigned long long)slot);
        fpt_hip_batch_destroy(b);
        return 1;
    }
    b->total_slots = uint32_t(slot);
    b->pixels = size_t(off);
    for (const FptRenderConfig &c : b->host) b->lowered = b->lowered || fpt::has_lowered_tables(c);
    hipError_t e = hipMalloc(&b->dcfgs, sizeof(FptRenderConfig) * count);
    if (e == hipSuccess) e = hipMalloc(&b->dscenes, sizeof(fpt::BatchScene) * count);
    if (e == hipSuccess) e = hipMalloc(&b->counter, sizeof(uint32_t));
    if (e == hipSuccess)
        e = hipMemcpy(b->dcfgs, b->host.data(), sizeof(FptRenderConfig) * count, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(b->dscenes, b->scenes.data(), sizeof(fpt::BatchScene) * count, hipMemcpyHostToDevice);
    // Focus prepass of every scene with depth of field and no focus distance
    // (device configs only: the host copies keep the caller's values).
    bool need_focus = false;
    for (uint32_t i = 0; i < count; ++i)
        need_focus = need_focus || (configs[i].camera_dof > 0.0f && !(configs[i].focus_distance > 0.0f));
    if (e == hipSuccess && need_focus) {
        e = fpt::launch_focus_prepass(b->dcfgs, count, nullptr);
        if (e == hipSuccess) e = hipDeviceSynchronize();
    }
    if (e != hipSuccess) {
        set_error(error, error_len, "batch upload failed: %s", hipGetErrorString(e));
        fpt_hip_batch_destroy(b);
        return 1;
    }
    // Longest-first hand-out of the batch's tiles, over all scenes.
    {
        const uint32_t ntiles = b->total_slots / 64u;
        std::vector<float> h(ntiles, 0.0f);
        for (uint32_t i = 0; i < count && e == hipSuccess; ++i)
            e = scene_tile_costs(b->dcfgs + i, b->host[i], h.data() + b->scenes[i].slot_begin / 64u);
        const std::vector<uint32_t> order = tile_order_from_costs(h);
        if (e == hipSuccess) e = hipMalloc(&b->tile_order, size_t(ntiles) * sizeof(uint32_t));
        if (e == hipSuccess)
            e = hipMemcpy(b->tile_order, order.data(), size_t(ntiles) * sizeof(uint32_t), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            set_error(error, error_len, "batch tile order failed: %s", hipGetErrorString(e));
            fpt_hip_batch_destroy(b);
            return 1;
        }
    }
    *out = b;
    return 0;
}

size_t fpt_hip_batch_pixels(const FptHipBatch *batch) { return batch ? batch->pixels : 0; }

int fpt_hip_batch_tune(FptHipBatch *batch, uint32_t workgroups, uint32_t event_batch, uint32_t use_graph, char *error,
                       size_t error_len) {
    if (!batch) {
        set_error(error, error_len, "missing batch");
        return 1;
    }
    if (event_batch > 128u || workgroups > 65536u) {
        set_error(error, error_len, "invalid launch shape (workgroups %u, event_batch %u)", workgroups, event_batch);
        return 1;
    }
    batch->workgroups = workgroups;
    batch->event_batch = event_batch;
    batch->use_graph = use_graph;
    return 0;
}

int fpt_hip_batch_render(FptHipBatch *batch, float *device_accum, void *stream, char *error, size_t error_len) {
    if (!batch || !device_accum) {
        set_error(error, error_len, "missing batch or accumulation buffer");
        return 1;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint32_t wg = batch->workgroups ? batch->workgroups : default_workgroups();
    float4 *acc = reinterpret_cast<float4 *>(device_accum);
    if (!batch->use_graph) {
        HIP_TRY(fpt::launch_sdf_batch(batch->dcfgs, batch->dscenes, uint32_t(batch->host.size()), batch->total_slots, acc,
                                      batch->counter, wg, int(batch->event_batch), 0u, nullptr, batch->lowered,
                                      batch->tile_order_mode ? batch->tile_order : nullptr, s),
                "batch kernel launch");
        return 0;
    }
    if (!batch->exec || batch->graph_accum != device_accum || batch->graph_wg != wg ||
        batch->graph_eb != batch->event_batch || batch->graph_order != batch->tile_order_mode) {
        batch_drop_graph(batch);
        hipStream_t cap = nullptr;
        HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "capture stream");
        hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess)
            e = fpt::launch_sdf_batch(batch->dcfgs, batch->dscenes, uint32_t(batch->host.size()), batch->total_slots,
                                      acc, batch->counter, wg, int(batch->event_batch), 0u, nullptr, batch->lowered,
                                      batch->tile_order_mode ? batch->tile_order : nullptr, cap);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(cap, &g);
        (void)hipStreamDestroy(cap);
        if (e == hipSuccess) e = e2;
        if (e == hipSuccess) e = hipGraphInstantiate(&batch->exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            set_error(error, error_len, "batch graph capture failed: %s", hipGetErrorString(e));
            return 1;
        }
        batch->graph = g;
        batch->graph_accum = device_accum;
        batch->graph_wg = wg;
        batch->graph_eb = batch->event_batch;
        batch->graph_order = batch->tile_order_mode;
    }
    HIP_TRY(hipGraphLaunch(batch->exec, s), "batch graph launch");
    return 0;
}

void fpt_hip_batch_destroy(FptHipBatch *batch) {
    if (!batch) return;
    batch_drop_graph(batch);
    if (batch->dcfgs) (void)hipFree(batch->dcfgs);
    if (batch->dscenes) (void)hipFree(batch->dscenes);
    if (batch->counter) (void)hipFree(batch->counter);
    if (batch->tile_order) (void)hipFree(batch->tile_order);
    delete batch;
}

int fpt_hip_render_batch(const FptRenderConfig *configs, uint32_t count, float *linear_output, size_t linear_output_len,
                         double *elapsed_ms, char *error, size_t error_len) {
    FptHipBatch *b = nullptr;
    if (fpt_hip_batch_create(configs, count, &b, error, error_len)) return 1;
    struct G {
        FptHipBatch *b;
        ~G() { fpt_hip_batch_destroy(b); }
    } guard{b};
    if (linear_output && linear_output_len < b->pixels * 4) {
        set_error(error, error_len, "linear output buffer is too small (%zu < %zu)", linear_output_len, b->pixels * 4);
        return 1;
    }
    Stream st;
    HIP_TRY(hipStreamCreate(&st.s), "hipStreamCreate");
    DevBuf acc;
    HIP_TRY(acc.alloc(b->pixels * sizeof(float4)), "batch accumulation allocation");
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0), "hipEventCreate");
    HIP_TRY(hipEventCreate(&e1), "hipEventCreate");
    HIP_TRY(hipEventRecord(e0, st.s), "hipEventRecord");
    if (fpt_hip_batch_render(b, acc.as<float>(), st.s, error, error_len)) return 1;
    HIP_TRY(hipEventRecord(e1, st.s), "hipEventRecord");
    HIP_TRY(hipStreamSynchronize(st.s), "batch render");
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (elapsed_ms) *elapsed_ms = ms;
    if (linear_output)
        HIP_TRY(hipMemcpy(linear_output, acc.p, b->pixels * sizeof(float4), hipMemcpyDeviceToHost), "batch readback");
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free batched render: the batch's device state (scene configs,
// scene table, work counter) lives in one file-scope device array instead of
// hipMalloc'd buffers.  Layout of the array, 256-byte aligned sections:
//   [0, C)          FptRenderConfig[FPT_NOALLOC_MAX_SCENES]
//   [C, C + T)      fpt::BatchScene[FPT_NOALLOC_MAX_SCENES]
//   [C + T, + 256)  uint32_t work counter
// Every call on a device shares that device's copy and leaves its contents
// behind; the counter is cleared by the launch sequence's own fill kernel.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t align256(size_t n) { return (n + 255u) & ~size_t(255u); }
constexpr size_t kNoAllocScenesOffset = align256(sizeof(FptRenderConfig) * FPT_NOALLOC_MAX_SCENES);
constexpr size_t kNoAllocCounterOffset = align256(kNoAllocScenesOffset + sizeof(fpt::BatchScene) * FPT_NOALLOC_MAX_SCENES);
constexpr size_t kNoAllocScratchBytes = kNoAllocCounterOffset + 256u;
static_assert(kNoAllocScratchBytes <= 2u * 1024u * 1024u, "no-alloc scratch stays under 2 MiB");
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
// Device address of `scratch` on each device, taken once per device with that
// device current.  The array's name in host code is a host-side stand-in, not
// a device address: only these pointers reach kernels and copies.
std::mutex g_scratch_mu;
std::vector<unsigned char *> g_scratch_ptr;

int noalloc_scratch(unsigned char **out, char *error, size_t error_len) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (size_t(dev) >= g_scratch_ptr.size()) {
        int n = 0;
        HIP_TRY(hipGetDeviceCount(&n), "hipGetDeviceCount");
        g_scratch_ptr.resize(std::max(size_t(n), size_t(dev) + 1u), nullptr);
    }
    if (!g_scratch_ptr[size_t(dev)]) {
        void *ptr = nullptr;
        HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[size_t(dev)] = static_cast<unsigned char *>(ptr);
    }
    *out = g_scratch_ptr[size_t(dev)];
    return 0;
}
}  // namespace

extern "C" {

int fpt_hip_render_batch_noalloc(const FptRenderConfig *configs, uint32_t count, float *device_accum,
                                 size_t device_accum_len, void *stream, char *error, size_t error_len) {
    if (!configs || !device_accum) {
        set_error(error, error_len, "missing scene configs or accumulation buffer");
        return 1;
    }
    if (count == 0u || count > FPT_NOALLOC_MAX_SCENES) {
        set_error(error, error_len, "invalid scene count %u (the allocation-free batch holds 1..%u scenes)", count,
                  unsigned(FPT_NOALLOC_MAX_SCENES));
        return 1;
    }
    fpt::BatchScene scenes[FPT_NOALLOC_MAX_SCENES];
    uint64_t slot = 0, off = 0;
    bool lowered = false, need_focus = false;
    for (uint32_t i = 0; i < count; ++i) {
        const FptRenderConfig &c = configs[i];
        if (validate_config(&c, error, error_len)) return 1;
        if (reject_mandelbulber(&c, "batched rendering", error, error_len)) return 1;
        if (c.renderer_backend != FPT_RENDERER_SDF) {
            set_error(error, error_len, "scene %u: batched rendering supports the SDF backend only", i);
            return 1;
        }
        // Reading the HDRI file is the caller's step here
        // (fpt_hip_hydrate_hdri); a hydrated config renders as in a batch.
        if (c.world[0] == 2.0f && c.hdri_enabled != 0u && c.hdri_width == 0u) {
            set_error(error, error_len,
                      "scene %u: HDRI lighting needs a hydrated config (fpt_hip_hydrate_hdri) on the allocation-free path",
                      i);
            return 1;
        }
        scenes[i].slot_begin = uint32_t(slot);
        scenes[i].slots = ((c.width + 7u) / 8u) * ((c.height + 7u) / 8u) * 64u;
        scenes[i].width = c.width;
        scenes[i].height = c.height;
        scenes[i].accum_offset = off;
        slot += scenes[i].slots;
        off += uint64_t(c.width) * c.height;
        lowered = lowered || fpt::has_lowered_tables(c);
        need_focus = need_focus || (c.camera_dof > 0.0f && !(c.focus_distance > 0.0f));
    }
    if (slot > 0xffffffffull) {
        set_error(error, error_len, "batch has too many pixels (%llu slots)", (unsigned long long)slot);
        return 1;
    }
    if (device_accum_len < off * 4u) {
        set_error(error, error_len, "accumulation buffer is too small (%zu < %llu floats)", device_accum_len,
                  (unsigned long long)(off * 4u));
        return 1;
    }
    if (check_device(error, error_len)) return 1;
    unsigned char *base = nullptr;
    if (noalloc_scratch(&base, error, error_len)) return 1;
    FptRenderConfig *dcfgs = reinterpret_cast<FptRenderConfig *>(base);
    fpt::BatchScene *dscenes = reinterpret_cast<fpt::BatchScene *>(base + kNoAllocScenesOffset);
    uint32_t *counter = reinterpret_cast<uint32_t *>(base + kNoAllocCounterOffset);
    hipStream_t s = static_cast<hipStream_t>(stream);
    // Host sources are pageable, so both copies complete before they return.
    HIP_TRY(hipMemcpyAsync(dcfgs, configs, sizeof(FptRenderConfig) * count, hipMemcpyHostToDevice, s),
            "scene config upload");
    HIP_TRY(hipMemcpyAsync(dscenes, scenes, sizeof(fpt::BatchScene) * count, hipMemcpyHostToDevice, s),
            "scene table upload");
    if (need_focus) HIP_TRY(fpt::launch_focus_prepass(dcfgs, count, s), "focus-distance prepass");
    HIP_TRY(fpt::launch_sdf_batch(dcfgs, dscenes, count, uint32_t(slot), reinterpret_cast<float4 *>(device_accum),
                                  counter, default_workgroups(), 0, 0u, nullptr, lowered, nullptr, s),
            "batch kernel launch");
    return 0;
}

// ---------------------------------------------------------------------------
// Autotuner: candidate launch shapes timed on a proxy of the workload (the
// plan's first samples, up to 16 / the batch capped at one sample per pixel),
// probed in a rotated order every round so drift does not favour a position,
// median of the rounds, fastest median wins.
// ---------------------------------------------------------------------------
int fpt_hip_autotune(FptHipPlan *plan, FptHipBatch *batch, float *device_accum, void *stream, uint32_t rounds,
                     FptTuneResult *result, char *error, size_t error_len) {
    if ((!plan) == (!batch) || !device_accum) {
        set_error(error, error_len, "autotune needs exactly one of plan/batch and a device buffer");
        return 1;
    }
    if (plan && plan->host.renderer_backend != FPT_RENDERER_SDF) {
        set_error(error, error_len, "autotune applies to the SDF backend");
        return 1;
    }
    if (plan && reject_mandelbulber(&plan->host, "the persistent-kernel autotuner", error, error_len)) return 1;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint32_t cus = default_workgroups() / 2u;
    struct Shape {
        uint32_t workgroups, event_batch, order, groups;
    };
    std::vector<Shape> cand;
    // With the path stash a workgroup holds 53-57 KB of LDS, so at most two
    // are resident per compute unit, and the event batch counts the wave's
    // up to 128 paths.  A plan's event batch stays at the launch's default
    // (0): the best batch depends on whether the frame runs in sample runs
    // (64 with, 48 without), and a few-sample proxy does not split as the
    // frame does.
    const bool stash = plan ? fpt::persistent_uses_stash(plan->host, plan->proposer.enabled) : fpt::batch_uses_stash();
    const std::vector<uint32_t> per_cus = stash ? std::vector<uint32_t>{1u, 2u} : std::vector<uint32_t>{1u, 2u, 3u, 4u};
    const std::vector<uint32_t> ebs = !stash ? std::vector<uint32_t>{8u, 16u, 24u, 32u, 40u}
                                      : plan ? std::vector<uint32_t>{0u}
                                             : std::vector<uint32_t>{32u, 40u, 48u, 56u, 64u};
    const uint32_t default_eb = stash && plan ? 0u : uint32_t(fpt::default_event_batch(stash, true));
    // Sample runs against one slot per pixel, where the proxy is the whole
    // frame (at most 16 samples per pixel) and so measures the choice.
    const uint32_t spp = plan ? clamp_samples(plan->host.samples) : 0u;
    const uint32_t plan_groups = plan ? plan->sample_groups : 0u;
    // Candidates set the plan's run count while they are timed; an early
    // return (a failed launch) puts the caller's value back.
    struct GroupsGuard {
        FptHipPlan *p;
        uint32_t g;
        ~GroupsGuard() {
            if (p) p->sample_groups = g;
        }
    } groups_guard{plan, plan_groups};
    std::vector<uint32_t> groups{plan_groups};
    if (plan && stash && spp <= 16u && plan_groups > 1u) groups.push_back(1u);
    for (uint32_t g : groups)
        for (uint32_t order : {1u, 0u})
            for (uint32_t per_cu : per_cus)
                for (uint32_t eb : ebs) cand.push_back({per_cu * cus, eb, order, g});
    rounds = std::max<uint32_t>(1u, std::min<uint32_t>(rounds, 16u));
    std::vector<std::vector<float>> times(cand.size());
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0), "hipEventCreate");
    HIP_TRY(hipEventCreate(&e1), "hipEventCreate");
    float4 *acc = reinterpret_cast<float4 *>(device_accum);
    if (plan && !plan->counter) HIP_TRY(hipMalloc(&plan->counter, sizeof(uint32_t)), "counter allocation");
    const uint32_t plan_proxy_spp = plan ? std::min<uint32_t>(clamp_samples(plan->host.samples), 16u) : 1u;
    auto launch = [&](const Shape &sh) -> hipError_t {
        if (plan) {  // proxy: the frame's first samples (up to 16)
            fpt::SampleSplit split;
            plan->sample_groups = sh.groups;
            const hipError_t e = plan_split(plan, 0u, plan_proxy_spp, split);
            if (e != hipSuccess) return e;
            return fpt::launch_sdf_persistent(plan->dcfg, plan->host, acc, 0u, plan_proxy_spp, plan->counter,
                                              sh.workgroups, int(sh.event_batch), int(plan->refill), nullptr,
                                              plan->proposer.enabled ? &plan->proposer.dev : nullptr,
                                              sh.order ? plan->tile_order : nullptr, split, s);
        }
        return fpt::launch_sdf_batch(batch->dcfgs, batch->dscenes, uint32_t(batch->host.size()), batch->total_slots, acc,
                                     batch->counter, sh.workgroups, int(sh.event_batch), 1u, nullptr, batch->lowered,
                                     sh.order ? batch->tile_order : nullptr, s);
    };
    HIP_TRY(launch(cand[0]), "autotune warm-up");
    for (uint32_t r = 0; r < rounds; ++r) {
        for (size_t j = 0; j < cand.size(); ++j) {
            size_t c = (j + r * 7u) % cand.size();  // rotated probe order
            HIP_TRY(hipEventRecord(e0, s), "hipEventRecord");
            HIP_TRY(launch(cand[c]), "autotune launch");
            HIP_TRY(hipEventRecord(e1, s), "hipEventRecord");
            HIP_TRY(hipEventSynchronize(e1), "autotune");
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
            times[c].push_back(ms);
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    size_t best = 0;
    float best_ms = 1e30f, default_ms = 0.f;
    for (size_t c = 0; c < cand.size(); ++c) {
        std::vector<float> t = times[c];
        std::sort(t.begin(), t.end());
        float med = t[t.size() / 2];
        if (cand[c].workgroups == 2u * cus && cand[c].event_batch == default_eb && cand[c].order == 1u &&
            cand[c].groups == plan_groups)
            default_ms = med;
        if (med < best_ms) {
            best_ms = med;
            best = c;
        }
    }
    if (plan) {
        plan->kernel = 1u;
        plan->workgroups = cand[best].workgroups;
        plan->event_batch = cand[best].event_batch;
        plan->tile_order_mode = cand[best].order;
        groups_guard.g = cand[best].groups;
    } else {
        batch->workgroups = cand[best].workgroups;
        batch->event_batch = cand[best].event_batch;
        batch->tile_order_mode = cand[best].order;
    }
    if (result) {
        result->workgroups = cand[best].workgroups;
        result->event_batch = cand[best].event_batch;
        result->tile_order = cand[best].order;
        result->sample_groups = cand[best].groups;
        result->candidates = uint32_t(cand.size());
        result->rounds = rounds;
        result->best_ms = best_ms;
        result->default_ms = default_ms;
    }
    return 0;
}

}  // extern "C"

extern "C" {

int fpt_hip_set_tile_order(FptHipPlan *plan, FptHipBatch *batch, uint32_t mode, char *error, size_t error_len) {
    if ((!plan) == (!batch) || mode > 1u) {
        set_error(error, error_len, "tile order needs exactly one of plan/batch and mode 0 or 1");
        return 1;
    }
    if (plan) plan->tile_order_mode = mode;
    else batch->tile_order_mode = mode;
    return 0;
}

int fpt_metal_device_name(char *name, size_t name_len) {
    if (!name || name_len == 0) return 1;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        std::snprintf(name, name_len, "%s", "no HIP device");
        return 1;
    }
    std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    return 0;
}

int fpt_metal_render(const char *metallib_path, const char *stitch_metallib_path, const char *stitch_archive_path,
                     const char *output_path, const char *shader_source, size_t shader_source_len,
                     const FptRenderConfig *config, double *build_ms, double *elapsed_ms,
                     uint64_t *voxel_memory_bytes, uint32_t *voxel_active_bricks, uint64_t *voxel_active_cells,
                     uint32_t *voxel_rejected_bricks, FptBoundGridStats *bound_grid_stats,
                     uint32_t *stitch_cache_status, FptStitchValidationStats *stitch_validation_stats,
                     FptStitchPipelineStats *stitch_pipeline_stats, FptSdfProfileStats *sdf_profile_stats,
                     float *linear_output, size_t linear_output_len, char *error, size_t error_len) {
    (void)metallib_path; (void)stitch_metallib_path; (void)stitch_archive_path;
    (void)shader_source; (void)shader_source_len;
    if (validate_config(config, error, error_len)) return 1;
    if (build_ms) *build_ms = 0.0;
    if (elapsed_ms) *elapsed_ms = 0.0;
    if (voxel_memory_bytes) *voxel_memory_bytes = 0;
    if (voxel_active_bricks) *voxel_active_bricks = 0;
    if (voxel_active_cells) *voxel_active_cells = 0;
    if (voxel_rejected_bricks) *voxel_rejected_bricks = 0;
    if (bound_grid_stats) std::memset(bound_grid_stats, 0, sizeof(*bound_grid_stats));
    if (stitch_cache_status) *stitch_cache_status = 0;
    if (stitch_validation_stats) std::memset(stitch_validation_stats, 0, sizeof(*stitch_validation_stats));
    if (stitch_pipeline_stats) std::memset(stitch_pipeline_stats, 0, sizeof(*stitch_pipeline_stats));
    if (sdf_profile_stats) std::memset(sdf_profile_stats, 0, sizeof(*sdf_profile_stats));
    const size_t pixels = size_t(config->width) * config->height;
    if (linear_output && linear_output_len < pixels * 4) {  // metal_bridge.mm:4901-4906
        set_error(error, error_len, "linear output buffer is too small (%zu < %zu)", linear_output_len, pixels * 4);
        return 1;
    }
    FptHipPlan *plan = nullptr;
    if (fpt_hip_plan_create(config, &plan, error, error_len)) return 1;
    struct PlanGuard {
        FptHipPlan *p;
        ~PlanGuard() { fpt_hip_plan_destroy(p); }
    } guard{plan};
    if (build_ms) *build_ms = plan->build_ms;
    if (config->sdf_topology_specialization != 0u) {  // metal_bridge.mm:4469-4500: runtime-source render
        if (config->renderer_backend != FPT_RENDERER_SDF) {
            set_error(error, error_len, "runtime-source compilation requires the direct SDF renderer");
            return 1;
        }
        const int rc = plan_specialize(plan, error, error_len);
        const fpt::SpecializeReport &r = plan->spec_report;
        if (stitch_cache_status) *stitch_cache_status = r.cache_status;
        if (stitch_validation_stats) {
            stitch_validation_stats->sample_count = r.sample_count;
            stitch_validation_stats->distance_failures = r.distance_failures;
            stitch_validation_stats->gradient_failures = r.gradient_failures;
            stitch_validation_stats->max_distance_error = r.max_distance_error;
            stitch_validation_stats->max_gradient_error = r.max_gradient_error;
        }
        if (stitch_pipeline_stats) {
            FptStitchPipelineStats &ps = *stitch_pipeline_stats;
            const uint32_t n = std::min<uint32_t>(config->sdf_program_count, FPT_SDF_PROGRAM_MAX_OPS);
            ps.instruction_count = n;
            uint32_t prev = 0xffffffffu;
            for (uint32_t i = 0; i < n; ++i) {
                const FptSdfInstruction &op = config->sdf_program[i];
                const bool prim = op.opcode == FPT_SDF_OP_SPHERE || op.opcode == FPT_SDF_OP_BOX ||
                                  op.opcode == FPT_SDF_OP_PLANE || op.opcode == FPT_SDF_OP_MANDELBULB ||
                                  op.opcode == FPT_SDF_OP_MANDELBOX || op.opcode == FPT_SDF_OP_MENGER;
                if (prim) {
                    ps.primitive_instruction_count++;
                    if (op.opcode != prev) ps.primitive_type_runs++;
                    prev = op.opcode;
                    if (i) (op.flags == 1u ? ps.intersection_count : op.flags == 2u ? ps.subtraction_count
                                                                                      : ps.union_count)++;
                } else if (op.opcode != FPT_SDF_OP_ORBIT_ADD) {
                    ps.transform_instruction_count++;
                }
            }
            ps.thread_execution_width = 64u;
            ps.max_total_threads_per_threadgroup = 256u;
            ps.runtime_source_bytes = r.source_bytes;
            ps.runtime_library_compile_ms = r.compile_ms;
            ps.runtime_pipeline_link_ms = r.link_ms;
        }
        if (rc) return 1;
    }
    if (config->renderer_backend == FPT_RENDERER_VOXEL) {
        if (voxel_memory_bytes) *voxel_memory_bytes = plan->voxels.memory_bytes;
        if (voxel_active_bricks) *voxel_active_bricks = plan->voxels.active_bricks;
        if (voxel_rejected_bricks) *voxel_rejected_bricks = plan->voxels.rejected_bricks;
        if (voxel_active_cells) *voxel_active_cells = plan->voxels.active_cells;
    }
    Stream st;
    HIP_TRY(hipStreamCreate(&st.s), "hipStreamCreate");
    DevBuf accum;
    HIP_TRY(accum.alloc(pixels * sizeof(float4)), "accumulation buffer allocation");
    hipEvent_t ev0, ev1;
    HIP_TRY(hipEventCreate(&ev0), "hipEventCreate");
    HIP_TRY(hipEventCreate(&ev1), "hipEventCreate");
    HIP_TRY(hipEventRecord(ev0, st.s), "hipEventRecord");
    if (fpt_hip_plan_render(plan, accum.as<float>(), st.s, error, error_len)) return 1;
    HIP_TRY(hipEventRecord(ev1, st.s), "hipEventRecord");
    HIP_TRY(hipStreamSynchronize(st.s), "render");
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1), "hipEventElapsedTime");
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    if (elapsed_ms) *elapsed_ms = ms;
    if (linear_output) {
        HIP_TRY(hipMemcpy(linear_output, accum.p, pixels * sizeof(float4), hipMemcpyDeviceToHost), "linear readback");
    }
    if (sdf_profile_stats && config->renderer_backend == FPT_RENDERER_SDF) {
        // Counting pass (untimed), every pixel at stride 1.
        DevBuf prof, acc2;
        HIP_TRY(prof.alloc(sizeof(fpt::DeviceProfile)), "profile buffer");
        HIP_TRY(acc2.alloc(pixels * sizeof(float4)), "profile accumulation buffer");
        HIP_TRY(hipMemsetAsync(prof.p, 0, sizeof(fpt::DeviceProfile), st.s), "profile clear");
        const uint32_t spp = clamp_samples(config->samples), chunk = sdf_launch_chunk(plan);
        for (uint32_t s0 = 0; s0 < spp; s0 += chunk)
            HIP_TRY(launch_sdf(plan, acc2.as<float4>(), s0, chunk, prof.as<fpt::DeviceProfile>(), st.s),
                    "profile kernel launch");
        fpt::DeviceProfile hp;
        HIP_TRY(hipMemcpyAsync(&hp, prof.p, sizeof(hp), hipMemcpyDeviceToHost, st.s), "profile readback");
        HIP_TRY(hipStreamSynchronize(st.s), "profile pass");
        sdf_profile_stats->primary_steps = hp.primary_steps;
        sdf_profile_stats->secondary_steps = hp.secondary_steps;
        sdf_profile_stats->shadow_steps = hp.shadow_steps;
        sdf_profile_stats->normal_evals = hp.normal_evals;
        sdf_profile_stats->bounces = hp.bounces;
        sdf_profile_stats->pixels = hp.pixels;
        sdf_profile_stats->distance_evals = hp.distance_evals;
        sdf_profile_stats->march_orbit_iterations = hp.orbit_iterations;
    }
    if (output_path && output_path[0]) {
        DevBuf rgba_d;
        HIP_TRY(rgba_d.alloc(pixels * 4), "present buffer");
        HIP_TRY(fpt::launch_present(plan->dcfg, *config, accum.as<float4>(), rgba_d.as<uchar4>(), st.s),
                "present kernel launch");
        std::vector<uint8_t> rgba(pixels * 4);
        HIP_TRY(hipMemcpyAsync(rgba.data(), rgba_d.p, rgba.size(), hipMemcpyDeviceToHost, st.s), "present readback");
        HIP_TRY(hipStreamSynchronize(st.s), "present");
        if (unique_colour_sample(rgba) < 2) {  // metal_bridge.mm:5852-5857
            set_error(error, error_len, "render output is blank or single-colour");
            return 1;
        }
        if (!write_png(output_path, config->width, config->height, rgba, error, error_len)) return 1;
    }
    return 0;
}

int fpt_hip_trace_primary(const FptRenderConfig *config, FptRayHit *hits, size_t hit_count, char *error,
                          size_t error_len) {
    if (validate_config(config, error, error_len)) return 1;
    if (config->renderer_backend != FPT_RENDERER_SDF) {
        set_error(error, error_len, "fpt_hip_trace_primary needs the SDF backend");
        return 1;
    }
    const size_t pixels = size_t(config->width) * config->height;
    if (!hits || hit_count < pixels) {
        set_error(error, error_len, "hit buffer is too small (%zu < %zu)", hit_count, pixels);
        return 1;
    }
    if (check_device(error, error_len)) return 1;
    Stream st;
    HIP_TRY(hipStreamCreate(&st.s), "hipStreamCreate");
    DevBuf dcfg, dh;
    HIP_TRY(dcfg.alloc(sizeof(FptRenderConfig)), "config allocation");
    HIP_TRY(dh.alloc(pixels * sizeof(FptRayHit)), "hit buffer allocation");
    HIP_TRY(hipMemcpyAsync(dcfg.p, config, sizeof(FptRenderConfig), hipMemcpyHostToDevice, st.s), "config upload");
    HIP_TRY(fpt::launch_trace_primary(dcfg.as<FptRenderConfig>(), *config, dh.as<FptRayHit>(), st.s),
            "trace kernel launch");
    HIP_TRY(hipMemcpyAsync(hits, dh.p, pixels * sizeof(FptRayHit), hipMemcpyDeviceToHost, st.s), "hit readback");
    HIP_TRY(hipStreamSynchronize(st.s), "trace");
    return 0;
}

int fpt_mandelbulber_sample_field(const char *metallib_path, const char *shader_source, size_t shader_source_len,
                                  const FptRenderConfig *config, const float *points_xyzw, size_t point_count,
                                  FptMandelbulberFieldSample *samples, char *error, size_t error_len) {
    (void)metallib_path; (void)shader_source; (void)shader_source_len;
    if (!config || !points_xyzw || !samples) {
        set_error(error, error_len, "missing config, points or sample buffer");
        return 1;
    }
    if (point_count == 0) return 0;
    if (check_device(error, error_len)) return 1;
    Stream st;
    HIP_TRY(hipStreamCreate(&st.s), "hipStreamCreate");
    DevBuf dcfg, dp, ds;
    HIP_TRY(dcfg.alloc(sizeof(FptRenderConfig)), "config allocation");
    HIP_TRY(dp.alloc(point_count * 16), "point buffer allocation");
    HIP_TRY(ds.alloc(point_count * sizeof(FptMandelbulberFieldSample)), "sample buffer allocation");
    HIP_TRY(hipMemcpyAsync(dcfg.p, config, sizeof(FptRenderConfig), hipMemcpyHostToDevice, st.s), "config upload");
    HIP_TRY(hipMemcpyAsync(dp.p, points_xyzw, point_count * 16, hipMemcpyHostToDevice, st.s), "point upload");
    HIP_TRY(fpt::launch_field_sample(dcfg.as<FptRenderConfig>(), dp.as<float4>(), point_count,
                                     ds.as<FptMandelbulberFieldSample>(), st.s),
            "field kernel launch");
    HIP_TRY(hipMemcpyAsync(samples, ds.p, point_count * sizeof(FptMandelbulberFieldSample), hipMemcpyDeviceToHost,
                           st.s),
            "sample readback");
    HIP_TRY(hipStreamSynchronize(st.s), "field sampling");
    return 0;
}

int fpt_hip_program_surface(const FptRenderConfig *config, const float *points_xyzw, size_t point_count,
                            float *out_dgrad, char *error, size_t error_len) {
    if (!config || !points_xyzw || !out_dgrad) {
        set_error(error, error_len, "missing config, points or output buffer");
        return 1;
    }
    if (point_count == 0) return 0;
    if (check_device(error, error_len)) return 1;
    Stream st;
    HIP_TRY(hipStreamCreate(&st.s), "hipStreamCreate");
    DevBuf dcfg, dp, dout;
    HIP_TRY(dcfg.alloc(sizeof(FptRenderConfig)), "config allocation");
    HIP_TRY(dp.alloc(point_count * 16), "point buffer allocation");
    HIP_TRY(dout.alloc(point_count * 16), "output buffer allocation");
    HIP_TRY(hipMemcpyAsync(dcfg.p, config, sizeof(FptRenderConfig), hipMemcpyHostToDevice, st.s), "config upload");
    HIP_TRY(hipMemcpyAsync(dp.p, points_xyzw, point_count * 16, hipMemcpyHostToDevice, st.s), "point upload");
    HIP_TRY(fpt::launch_surface_sample(dcfg.as<FptRenderConfig>(), dp.as<float4>(), point_count, dout.as<float4>(),
                                       st.s),
            "surface kernel launch");
    HIP_TRY(hipMemcpyAsync(out_dgrad, dout.p, point_count * 16, hipMemcpyDeviceToHost, st.s), "readback");
    HIP_TRY(hipStreamSynchronize(st.s), "surface sampling");
    return 0;
}

int fpt_hip_voxel_build(const FptRenderConfig *config, uint32_t *page_table, size_t page_table_len, void *records,
                        size_t records_len, uint16_t *distance, size_t distance_len, uint8_t *material,
                        size_t material_len, float *palette, size_t palette_len, FptVoxelBuildInfo *info,
                        char *error, size_t error_len) {
    if (!config) {
        set_error(error, error_len, "missing render config");
        return 1;
    }
    if (reject_mandelbulber(config, "the voxel builder", error, error_len)) return 1;
    if (check_device(error, error_len)) return 1;
    Stream st;
    HIP_TRY(hipStreamCreate(&st.s), "hipStreamCreate");
    DevBuf dcfg;
    HIP_TRY(dcfg.alloc(sizeof(FptRenderConfig)), "config allocation");
    HIP_TRY(hipMemcpyAsync(dcfg.p, config, sizeof(FptRenderConfig), hipMemcpyHostToDevice, st.s), "config upload");
    fpt::VoxelStore vs;
    struct Guard {
        fpt::VoxelStore *v;
        ~Guard() { fpt::voxel_free(*v); }
    } guard{&vs};
    auto t0 = std::chrono::steady_clock::now();
    if (fpt::voxel_build(*config, dcfg.as<FptRenderConfig>(), vs, st.s, error, error_len)) {
        g_last_error = error && error_len ? error : "voxel build failed";
        return 1;
    }
    double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    FptVoxelBuildInfo inf = {vs.storage, vs.brick_size,   vs.bricks_per_axis, vs.active_bricks, vs.active_cells,
                             vs.memory_bytes, ms,         vs.rejected_bricks, 0u};
    if (info) *info = inf;
    const size_t G = vs.bricks_per_axis;
    const size_t res = config->voxel_resolution;
    const size_t cpb = size_t(vs.brick_size) * vs.brick_size * vs.brick_size;
    const size_t pages = vs.active_bricks;
    auto fetch = [&](void *dst, size_t have, size_t need, const void *src, size_t elem, const char *what) -> int {
        if (!dst) return 0;
        if (have < need) {
            set_error(error, error_len, "%s buffer is too small (%zu < %zu)", what, have, need);
            return 1;
        }
        if (need && src) {
            hipError_t e = hipMemcpyAsync(dst, src, need * elem, hipMemcpyDeviceToHost, st.s);
            if (e != hipSuccess) {
                set_error(error, error_len, "%s readback failed: %s", what, hipGetErrorString(e));
                return 1;
            }
        }
        return 0;
    };
    if (vs.storage == FPT_VOXEL_STORAGE_DENSE) {
        if (fetch(records, records_len, res * res * res, vs.pool, 12, "records")) return 1;
    } else {
        if (fetch(page_table, page_table_len, G * G * G, vs.page_table, 4, "page table")) return 1;
        if (vs.storage == FPT_VOXEL_STORAGE_SPARSE_BRICKS) {
            if (fetch(records, records_len, pages * cpb, vs.pool, 12, "records")) return 1;
        } else {
            if (fetch(distance, distance_len, pages * cpb, vs.distance, 2, "distance")) return 1;
            if (fetch(material, material_len, pages * cpb, vs.material, 1, "material")) return 1;
            if (fetch(palette, palette_len, 256 * 8, vs.palette, 4, "palette")) return 1;
        }
    }
    HIP_TRY(hipStreamSynchronize(st.s), "voxel readback");
    return 0;
}

int fpt_hip_voxel_trace_cells(const FptRenderConfig *config, const float *rays, size_t ray_count, uint32_t cap,
                              int32_t *cells, uint32_t *counts, float *hits, char *error, size_t error_len) {
    if (!config || !rays || !counts || !hits || (cap && !cells)) {
        set_error(error, error_len, "missing config, rays or output buffers");
        return 1;
    }
    if (ray_count == 0) return 0;
    if (check_device(error, error_len)) return 1;
    Stream st;
    HIP_TRY(hipStreamCreate(&st.s), "hipStreamCreate");
    DevBuf dcfg, dr, dc, dn, dh;
    HIP_TRY(dcfg.alloc(sizeof(FptRenderConfig)), "config allocation");
    HIP_TRY(hipMemcpyAsync(dcfg.p, config, sizeof(FptRenderConfig), hipMemcpyHostToDevice, st.s), "config upload");
    fpt::VoxelStore vs;
    struct Guard {
        fpt::VoxelStore *v;
        ~Guard() { fpt::voxel_free(*v); }
    } guard{&vs};
    if (fpt::voxel_build(*config, dcfg.as<FptRenderConfig>(), vs, st.s, error, error_len)) {
        g_last_error = error && error_len ? error : "voxel build failed";
        return 1;
    }
    HIP_TRY(dr.alloc(ray_count * 6 * sizeof(float)), "ray buffer allocation");
    HIP_TRY(dc.alloc(ray_count * size_t(cap) * 3 * sizeof(int32_t)), "cell buffer allocation");
    HIP_TRY(dn.alloc(ray_count * sizeof(uint32_t)), "count buffer allocation");
    HIP_TRY(dh.alloc(ray_count * 8 * sizeof(float)), "hit buffer allocation");
    HIP_TRY(hipMemcpyAsync(dr.p, rays, ray_count * 6 * sizeof(float), hipMemcpyHostToDevice, st.s), "ray upload");
    HIP_TRY(hipMemsetAsync(dc.p, 0, ray_count * size_t(cap) * 3 * sizeof(int32_t), st.s), "cell buffer clear");
    HIP_TRY(fpt::launch_voxel_trace_cells(dcfg.as<FptRenderConfig>(), vs, dr.as<float>(), ray_count, cap,
                                          dc.as<int>(), dn.as<uint32_t>(), dh.as<float>(), st.s),
            "trace kernel launch");
    if (cap)
        HIP_TRY(hipMemcpyAsync(cells, dc.p, ray_count * size_t(cap) * 3 * sizeof(int32_t), hipMemcpyDeviceToHost, st.s),
                "cell readback");
    HIP_TRY(hipMemcpyAsync(counts, dn.p, ray_count * sizeof(uint32_t), hipMemcpyDeviceToHost, st.s), "count readback");
    HIP_TRY(hipMemcpyAsync(hits, dh.p, ray_count * 8 * sizeof(float), hipMemcpyDeviceToHost, st.s), "hit readback");
    HIP_TRY(hipStreamSynchronize(st.s), "voxel trace");
    return 0;
}

}  // extern "C"

namespace {
// Diagnostic image of `config` into host rgba (width*height*4, PNG row
// order), on its own stream; elapsed = GPU time of the diagnostic kernel (and
// voxel build for the voxel backend), metal_bridge.mm:5865-6011.
int diagnostic_image(const FptRenderConfig *config, const FptDiagnosticConfig *diagnostic, uint8_t *rgba,
                     double *elapsed_ms, char *error, size_t error_len) {
    if (!config || !diagnostic) {
        set_error(error, error_len, "missing diagnostic render config");
        return 1;
    }
    if (validate_config(config, error, error_len)) return 1;
    if (diagnostic->mode > FPT_DIAGNOSTIC_SDF_BOUNCE_CONTRIBUTION) {
        set_error(error, error_len, "unknown diagnostic mode %u", diagnostic->mode);
        return 1;
    }
    const auto t0 = std::chrono::steady_clock::now();
    FptHipPlan *plan = nullptr;
    if (fpt_hip_plan_create(config, &plan, error, error_len)) return 1;
    struct PlanGuard {
        FptHipPlan *p;
        ~PlanGuard() { fpt_hip_plan_destroy(p); }
    } guard{plan};
    const size_t pixels = size_t(config->width) * config->height;
    Stream st;
    HIP_TRY(hipStreamCreate(&st.s), "hipStreamCreate");
    DevBuf out;
    HIP_TRY(out.alloc(pixels * 4), "diagnostic buffer allocation");
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0), "hipEventCreate");
    HIP_TRY(hipEventCreate(&e1), "hipEventCreate");
    HIP_TRY(hipEventRecord(e0, st.s), "hipEventRecord");
    hipError_t e = config->renderer_backend == FPT_RENDERER_VOXEL
                       ? fpt::launch_voxel_diagnostic(plan->dcfg, *config, plan->voxels, *diagnostic,
                                                      out.as<uchar4>(), st.s)
                       : fpt::launch_sdf_diagnostic(plan->dcfg, *config, *diagnostic, out.as<uchar4>(), st.s);
    HIP_TRY(e, "diagnostic kernel launch");
    HIP_TRY(hipEventRecord(e1, st.s), "hipEventRecord");
    HIP_TRY(hipMemcpyAsync(rgba, out.p, pixels * 4, hipMemcpyDeviceToHost, st.s), "diagnostic readback");
    HIP_TRY(hipStreamSynchronize(st.s), "diagnostic render");
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (elapsed_ms) {
        const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        *elapsed_ms = ms > 0.f ? double(ms) + plan->build_ms : wall;
    }
    return 0;
}
}  // namespace

extern "C" {

int fpt_metal_diagnostic_render(const char *metallib_path, const char *output_path, const char *shader_source,
                                size_t shader_source_len, const FptRenderConfig *config,
                                const FptDiagnosticConfig *diagnostic, double *elapsed_ms, char *error,
                                size_t error_len) {
    (void)metallib_path; (void)shader_source; (void)shader_source_len;
    if (!config || !diagnostic) {
        set_error(error, error_len, "missing diagnostic render config");
        return 1;
    }
    if (!output_path || !output_path[0]) {
        set_error(error, error_len, "missing diagnostic output path");
        return 1;
    }
    if (config->width == 0 || config->height == 0 || config->width > 16384 || config->height > 16384) {
        set_error(error, error_len, "invalid render size %ux%u", config->width, config->height);
        return 1;
    }
    std::vector<uint8_t> rgba(size_t(config->width) * config->height * 4);
    if (diagnostic_image(config, diagnostic, rgba.data(), elapsed_ms, error, error_len)) return 1;
    return write_png(output_path, config->width, config->height, rgba, error, error_len) ? 0 : 1;
}

int fpt_hip_diagnostic_render(const FptRenderConfig *config, const FptDiagnosticConfig *diagnostic, uint8_t *rgba,
                              size_t rgba_len, double *elapsed_ms, char *error, size_t error_len) {
    if (!config || !rgba) {
        set_error(error, error_len, "missing config or output buffer");
        return 1;
    }
    const size_t need = size_t(config->width) * config->height * 4;
    if (rgba_len < need) {
        set_error(error, error_len, "diagnostic buffer is too small (%zu < %zu)", rgba_len, need);
        return 1;
    }
    return diagnostic_image(config, diagnostic, rgba, elapsed_ms, error, error_len);
}

int fpt_compare_images(const char *baseline_path, const char *candidate_path, const char *report_path, char *error,
                       size_t error_len) {
    if (!report_path || !report_path[0]) {
        set_error(error, error_len, "missing comparison report path");
        return 1;
    }
    Image a, b;
    if (!read_png(baseline_path, a, error, error_len) || !read_png(candidate_path, b, error, error_len)) return 1;
    if (a.width != b.width || a.height != b.height) {
        set_error(error, error_len, "dimension mismatch: baseline %ux%u candidate %ux%u", a.width, a.height, b.width,
                  b.height);
        return 1;
    }
    // metal_bridge.mm:6077-6104: RGB channel statistics.
    const size_t pixels = size_t(a.width) * a.height;
    double abs_sum = 0, sq_sum = 0;
    uint32_t max_abs = 0;
    uint64_t changed = 0;
    for (size_t p = 0; p < pixels; ++p) {
        bool ch_changed = false;
        for (int ch = 0; ch < 3; ++ch) {
            const uint32_t d = uint32_t(std::abs(int(a.rgba[p * 4 + ch]) - int(b.rgba[p * 4 + ch])));
            abs_sum += d;
            sq_sum += double(d) * d;
            max_abs = std::max(max_abs, d);
            ch_changed |= d > 0;
        }
        changed += ch_changed;
    }
    const double samples = double(pixels * 3);
    const double mae = abs_sum / samples, rmse = std::sqrt(sq_sum / samples);
    const double changed_pct = pixels ? 100.0 * double(changed) / double(pixels) : 0.0;
    const size_t ua = unique_colour_sample(a.rgba), ub = unique_colour_sample(b.rgba);
    const CompareGate gate = gate_for(candidate_path);
    const double ssim = block_ssim(a, b), lf_ssim = low_frequency_ssim(a, b);
    const bool recognition = ub >= 16 && mae < gate.mae;
    const bool strict = ub >= 64 && mae < gate.strict_mae && lf_ssim >= gate.strict_ssim;
    std::string sheet_path(report_path);
    const size_t slash = sheet_path.find_last_of('/'), dot = sheet_path.find_last_of('.');
    if (dot != std::string::npos && (slash == std::string::npos || dot > slash)) sheet_path.resize(dot);
    sheet_path += ".comparison.png";
    uint32_t sw = 0, sh = 0;
    const std::vector<uint8_t> sheet = comparison_sheet(a, b, sw, sh);
    if (!write_png(sheet_path.c_str(), sw, sh, sheet, error, error_len)) return 1;
    FILE *f = std::fopen(report_path, "w");
    if (!f) {
        set_error(error, error_len, "failed to open report %s", report_path);
        return 1;
    }
    std::fprintf(f,
                 "{\n  \"baseline\": \"%s\",\n  \"candidate\": \"%s\",\n  \"width\": %u,\n  \"height\": %u,\n"
                 "  \"mean_absolute_error\": %.17g,\n  \"root_mean_square_error\": %.17g,\n"
                 "  \"luminance_ssim\": %.17g,\n  \"low_frequency_luminance_ssim\": %.17g,\n"
                 "  \"max_absolute_error\": %u,\n  \"changed_pixel_percent\": %.17g,\n  \"gate_class\": \"%s\",\n"
                 "  \"gate_mae_threshold\": %g,\n  \"strict_mae_threshold\": %g,\n  \"strict_ssim_threshold\": %g,\n"
                 "  \"comparison_sheet\": \"%s\",\n  \"unique_colours\": {\n    \"baseline\": %zu,\n"
                 "    \"candidate\": %zu\n  },\n  \"recognition_gate\": %s,\n  \"strict_gate\": %s\n}\n",
                 baseline_path, candidate_path, a.width, a.height, mae, rmse, ssim, lf_ssim, max_abs, changed_pct,
                 gate.name, gate.mae, gate.strict_mae, gate.strict_ssim, sheet_path.c_str(), ua, ub,
                 recognition ? "true" : "false", strict ? "true" : "false");
    std::fclose(f);
    return 0;
}

int fpt_test_async_stitch_context(const char *, const char *, const char *, const FptRenderConfig *,
                                  FptAsyncJitStats *, char *error, size_t error_len) {
    set_error(error, error_len, "Metal function stitching has no counterpart in this library");
    return 1;
}

// metal_bridge.mm:4116-4203: validates a lowered typed structure-of-arrays
// table against the interpreted program at 2^20 hashed points; fails when
// any point misses a tolerance.  The library path is ignored.
int fpt_test_typed_soa(const char *, const FptRenderConfig *config, FptStitchValidationStats *stats, char *error,
                       size_t error_len) {
    if (!config || !stats) {
        set_error(error, error_len, "missing typed-SoA validation input");
        return 1;
    }
    std::memset(stats, 0, sizeof(*stats));
    const FptTypedSoAProgram &t = config->sdf_typed_soa;
    if (config->sdf_id != FPT_SDF_PROGRAM || t.sphere_count + t.box_count + t.plane_count == 0u) {
        set_error(error, error_len, "typed-SoA validation requires a lowered typed program");
        return 1;
    }
    if (check_device(error, error_len)) return 1;
    constexpr uint32_t kSamples = 1u << 20;
    Stream st;
    HIP_TRY(hipStreamCreate(&st.s), "hipStreamCreate");
    DevBuf dcfg, dcounts;
    HIP_TRY(dcfg.alloc(sizeof(FptRenderConfig)), "config allocation");
    HIP_TRY(dcounts.alloc(4 * sizeof(uint32_t)), "counter allocation");
    HIP_TRY(hipMemcpyAsync(dcfg.p, config, sizeof(FptRenderConfig), hipMemcpyHostToDevice, st.s), "config upload");
    HIP_TRY(hipMemsetAsync(dcounts.p, 0, 4 * sizeof(uint32_t), st.s), "counter reset");
    HIP_TRY(fpt::launch_typed_soa_validation(dcfg.as<FptRenderConfig>(), kSamples, dcounts.as<uint32_t>(), st.s),
            "typed-SoA validation launch");
    uint32_t counts[4] = {};
    HIP_TRY(hipMemcpyAsync(counts, dcounts.p, sizeof(counts), hipMemcpyDeviceToHost, st.s), "readback");
    HIP_TRY(hipStreamSynchronize(st.s), "typed-SoA validation");
    stats->sample_count = kSamples;
    stats->distance_failures = counts[0];
    stats->gradient_failures = counts[1];
    std::memcpy(&stats->max_distance_error, &counts[2], sizeof(float));
    std::memcpy(&stats->max_gradient_error, &counts[3], sizeof(float));
    if (counts[0] != 0u || counts[1] != 0u) {
        set_error(error, error_len, "typed-SoA validation failed: %u distance, %u gradient", counts[0], counts[1]);
        return 1;
    }
    return 0;
}

int fpt_metal_preview(const char *, const char *, const char *const *, const char *, size_t,
                      const FptRenderConfig *, const FptRenderConfig *, uint32_t, uint32_t, char *error,
                      size_t error_len) {
    set_error(error, error_len, "the interactive preview has no counterpart in this library");
    return 1;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Device plumbing for callers without their own HIP runtime handle (tests,
// bench): buffers, streams, a graph-safe fill, and HIP-event timing of
// repeated renders on the launch stream.
// ---------------------------------------------------------------------------
namespace {
__global__ void fill_f32_kernel(float *dst, float v, size_t n) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    size_t stride = size_t(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) dst[i] = v;
}

// Device-to-device copy as a kernel (16-byte elements, grid-stride), so its
// traffic shows up in the kernel's own counters.
__global__ void copy_u4_kernel(uint4 *__restrict__ dst, const uint4 *__restrict__ src, size_t n) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    size_t stride = size_t(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}
}  // namespace

extern "C" {

int fpt_hip_device_alloc(size_t bytes, void **ptr, char *error, size_t error_len) {
    if (!ptr) {
        set_error(error, error_len, "missing pointer out-parameter");
        return 1;
    }
    *ptr = nullptr;
    if (check_device(error, error_len)) return 1;
    HIP_TRY(hipMalloc(ptr, bytes ? bytes : 1), "device allocation");
    return 0;
}

void fpt_hip_device_free(void *ptr) {
    if (ptr) (void)hipFree(ptr);
}

int fpt_hip_copy(void *dst, const void *src, size_t bytes, uint32_t to_device, char *error, size_t error_len) {
    if (to_device == 2u) {
        if (bytes % 16u != 0u || (reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) % 16u != 0u) {
            set_error(error, error_len, "device copies take 16-byte aligned pointers and sizes");
            return 1;
        }
        const size_t n = bytes / 16u;
        if (!n) return 0;
        const unsigned blocks = unsigned(std::min<size_t>((n + 255) / 256, 8192));
        hipLaunchKernelGGL(copy_u4_kernel, dim3(blocks), dim3(256), 0, nullptr, static_cast<uint4 *>(dst),
                           static_cast<const uint4 *>(src), n);
        HIP_TRY(hipGetLastError(), "copy kernel launch");
        HIP_TRY(hipDeviceSynchronize(), "device copy");
        return 0;
    }
    HIP_TRY(hipMemcpy(dst, src, bytes, to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost), "copy");
    return 0;
}

int fpt_hip_fill_f32(float *dst, float value, size_t count, void *stream, char *error, size_t error_len) {
    if (!count) return 0;
    unsigned blocks = unsigned(std::min<size_t>((count + 255) / 256, 65536));
    hipLaunchKernelGGL(fill_f32_kernel, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream), dst, value, count);
    HIP_TRY(hipGetLastError(), "fill kernel launch");
    return 0;
}

int fpt_hip_stream_create(void **stream, char *error, size_t error_len) {
    hipStream_t s = nullptr;
    if (check_device(error, error_len)) return 1;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    *stream = s;
    return 0;
}

void fpt_hip_stream_destroy(void *stream) {
    if (stream) (void)hipStreamDestroy(static_cast<hipStream_t>(stream));
}

int fpt_hip_synchronize(void *stream, char *error, size_t error_len) {
    HIP_TRY(stream ? hipStreamSynchronize(static_cast<hipStream_t>(stream)) : hipDeviceSynchronize(), "synchronize");
    return 0;
}

int fpt_hip_time_renders(FptHipPlan *plan, FptHipBatch *batch, float *device_accum, void *stream, uint32_t steps,
                         double *elapsed_ms, char *error, size_t error_len) {
    if ((!plan) == (!batch) || !device_accum || !elapsed_ms) {
        set_error(error, error_len, "time_renders needs exactly one of plan/batch, a buffer and an output");
        return 1;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipEvent_t e0, e1;
    HIP_TRY(hipStreamSynchronize(s), "synchronize");
    HIP_TRY(hipEventCreate(&e0), "hipEventCreate");
    HIP_TRY(hipEventCreate(&e1), "hipEventCreate");
    HIP_TRY(hipEventRecord(e0, s), "hipEventRecord");
    for (uint32_t i = 0; i < steps; ++i) {
        int rc = plan ? fpt_hip_plan_render(plan, device_accum, stream, error, error_len)
                      : fpt_hip_batch_render(batch, device_accum, stream, error, error_len);
        if (rc) return rc;
    }
    HIP_TRY(hipEventRecord(e1, s), "hipEventRecord");
    HIP_TRY(hipEventSynchronize(e1), "timed renders");
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *elapsed_ms = ms;
    return 0;
}

}  // extern "C"
