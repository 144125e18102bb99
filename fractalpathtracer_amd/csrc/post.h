// Tone mapping and colour grading of linear radiance (Shaders.metal:
// 6778-6825), shared by the present kernel and the diagnostic kernels.
#pragma once

#include "fpt_device.h"

namespace fpt {

__device__ __forceinline__ f3 gamma_srgb(f3 c) {  // SM:6778-6782
    auto one = [](float v) {
        float hi = 1.055f * powf(fmaxf(v, 0.0f), 1.0f / 2.4f) - 0.055f;
        return v <= 0.0031308f ? v * 12.92f : hi;
    };
    return mk(one(c.x), one(c.y), one(c.z));
}

__device__ __forceinline__ f3 aces(f3 x) {  // SM:6784-6791
    auto one = [](float v) {
        return clampf((v * (2.51f * v + 0.03f)) / (v * (2.43f * v + 0.59f) + 0.14f), 0.f, 1.f);
    };
    return mk(one(x.x), one(x.y), one(x.z));
}

__device__ __forceinline__ f3 pbr_neutral(f3 c) {  // SM:6793-6806
    const float start = 0.76f, desat = 0.15f;
    float x = fminf(c.x, fminf(c.y, c.z));
    float off = x < 0.08f ? x - 6.25f * x * x : 0.04f;
    c = c - mk(off, off, off);
    float peak = fmaxf(c.x, fmaxf(c.y, c.z));
    if (peak < start) return c;
    float d = 1.0f - start;
    float np = 1.0f - d * d / (peak + d - start);
    c = c * (np / peak);
    float g = 1.0f - 1.0f / (desat * (peak - np) + 1.0f);
    return vmix(c, mk(np, np, np), g);
}

__device__ __forceinline__ f3 tone_map(f3 c, const FptRenderConfig &cfg) {  // SM:6808-6816
    float n = cfg.post[0];
    if (n == 0.0f) return gamma_srgb(c);
    if (n == 1.0f) return mk(powf(c.x, 1.f / 2.4f), powf(c.y, 1.f / 2.4f), powf(c.z, 1.f / 2.4f));
    if (n == 2.0f) return mk(powf(c.x, 1.f / 2.6f), powf(c.y, 1.f / 2.6f), powf(c.z, 1.f / 2.6f));
    if (n == 3.0f) return aces(c);
    if (n == 4.0f) return pbr_neutral(c);
    return c;
}

__device__ __forceinline__ f3 post_process(f3 c, const FptRenderConfig &cfg) {  // SM:6818-6825
    c = tone_map(c, cfg);
    c = c * cfg.post[1] + mk(cfg.post[2], cfg.post[2], cfg.post[2]);
    c = (c - mk(0.5f, 0.5f, 0.5f)) * cfg.post[4] + mk(0.5f, 0.5f, 0.5f);
    float l = dot(c, mk(0.2126f, 0.7152f, 0.0722f));
    c = vmix(mk(l, l, l), c, cfg.post[3]);
    return mk(clampf(c.x, 0.f, 1.f), clampf(c.y, 0.f, 1.f), clampf(c.z, 0.f, 1.f));
}

}  // namespace fpt
