// Presentation: post-process the linear accumulation into RGBA8 with the
// reference's tone maps, exposure/contrast/saturation, chromatic aberration
// and highlight bloom, flipped vertically on write-out
// (Shaders.metal:6778-6825, 7462-7507, 8012-8025).
#include <hip/hip_runtime.h>

#include "fpt_device.h"
#include "fpt_internal.h"
#include "post.h"

namespace fpt {

__device__ __forceinline__ f3 px(const float4 *a, uint32_t x, uint32_t y, uint32_t w) {
    float4 v = a[size_t(y) * w + x];
    return mk(v.x, v.y, v.z);
}

__device__ f3 highlight_at(const float4 *acc, uint32_t x, uint32_t y, const FptRenderConfig &cfg) {  // SM:7466-7481
    float k = cfg.post[6];
    if (k <= 0.00001f) return mk(0.f, 0.f, 0.f);
    uint32_t w = cfg.width, h = cfg.height;
    uint32_t x0 = x > 0u ? x - 1u : x, x1 = x + 1u < w ? x + 1u : w - 1u;
    uint32_t y0 = y > 0u ? y - 1u : y, y1 = y + 1u < h ? y + 1u : h - 1u;
    f3 b = vmax0(px(acc, x, y, w) - mk(0.5f, 0.5f, 0.5f));
    b = b + px(acc, x1, y, w) * 0.25f + px(acc, x0, y, w) * 0.25f + px(acc, x, y1, w) * 0.25f +
        px(acc, x, y0, w) * 0.25f;
    return b * k;
}

__device__ f3 bilinear(const float4 *acc, float u, float v, const FptRenderConfig &cfg) {  // SM:7483-7496
    float fx = clampf(u, 0.f, 1.f) * float(cfg.width) - 0.5f;
    float fy = clampf(v, 0.f, 1.f) * float(cfg.height) - 0.5f;
    int bx = int(floorf(fx)), by = int(floorf(fy));
    float tx = fractf(fx), ty = fractf(fy);
    int wm = int(cfg.width) - 1, hm = int(cfg.height) - 1;
    uint32_t x0 = uint32_t(bx < 0 ? 0 : (bx > wm ? wm : bx)), y0 = uint32_t(by < 0 ? 0 : (by > hm ? hm : by));
    uint32_t x1 = x0 + 1u < cfg.width ? x0 + 1u : cfg.width - 1u;
    uint32_t y1 = y0 + 1u < cfg.height ? y0 + 1u : cfg.height - 1u;
    f3 c00 = px(acc, x0, y0, cfg.width), c10 = px(acc, x1, y0, cfg.width);
    f3 c01 = px(acc, x0, y1, cfg.width), c11 = px(acc, x1, y1, cfg.width);
    return vmix(vmix(c00, c10, tx), vmix(c01, c11, tx), ty);
}

__device__ f3 chromatic_at(const float4 *acc, uint32_t x, uint32_t y, const FptRenderConfig &cfg) {  // SM:7498-7507
    float u = (float(x) + 0.5f) / float(cfg.width), v = (float(y) + 0.5f) / float(cfg.height);
    float dx = u - 0.5f, dy = v - 0.5f;
    float len = sqrtf(dx * dx + dy * dy);
    float ox = dx * cfg.post[5] * len / 8.0f, oy = dy * cfg.post[5] * len / 8.0f;
    f3 centre = bilinear(acc, u, v, cfg);
    if (fabsf(cfg.post[5]) <= 0.00001f) return centre;
    f3 green = bilinear(acc, u - ox, v - oy, cfg);
    f3 blue = bilinear(acc, u - 2.0f * ox, v - 2.0f * oy, cfg);
    return mk(centre.x, green.y, blue.z);
}

__global__ __launch_bounds__(256) void present_kernel(const FptRenderConfig *__restrict__ cfgp,
                                                      const float4 *__restrict__ acc, uchar4 *__restrict__ out) {
    const FptRenderConfig &cfg = *cfgp;
    uint32_t x = blockIdx.x * 16u + (threadIdx.x & 15u);
    uint32_t y = blockIdx.y * 16u + (threadIdx.x >> 4);
    if (x >= cfg.width || y >= cfg.height) return;
    f3 c = post_process(chromatic_at(acc, x, y, cfg) + highlight_at(acc, x, y, cfg), cfg);
    out[size_t(cfg.height - 1u - y) * cfg.width + x] =
        make_uchar4((unsigned char)(clampf(c.x, 0.f, 1.f) * 255.0f), (unsigned char)(clampf(c.y, 0.f, 1.f) * 255.0f),
                    (unsigned char)(clampf(c.z, 0.f, 1.f) * 255.0f), 255);
}

hipError_t launch_present(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg, const float4 *accum,
                          uchar4 *out, hipStream_t stream) {
    dim3 grid((hcfg.width + 15u) / 16u, (hcfg.height + 15u) / 16u, 1u);
    hipLaunchKernelGGL(present_kernel, grid, dim3(256), 0, stream, dcfg, accum, out);
    return hipGetLastError();
}

}  // namespace fpt
