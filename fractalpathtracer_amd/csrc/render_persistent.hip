// Persistent, path-regenerating SDF path tracer for gfx950.
//
// The straightforward kernel (render.hip) runs one path per lane from camera
// to termination: a 64-lane wave then waits for its slowest ray at every
// march, and the per-hit work (analytic normal = forward-mode DE, material,
// sun setup) runs whenever any one lane reaches a hit.  Here every lane is a
// small state machine over the same path (Shaders.metal:5926-6026):
//
//   MARCH  one distance evaluation + advance per loop iteration
//          (Shaders.metal:3753-3783), for camera, bounce and shadow rays;
//   EVENT  the march has ended; the lane waits until enough lanes of its
//          wave are waiting (`event_batch`), then all of them run the hit /
//          shadow / scatter logic together, so its cost is paid once per
//          batch instead of once per lane;
//   IDLE   no pixel left.
//
// When a path ends the lane starts the pixel's next sample; when a pixel's
// samples are done it writes the running mean and takes the next pixel from
// a global counter (one atomic per wave, pixels in 8x8 tile order so a wave
// keeps working on neighbouring pixels).  Samples of a pixel are processed
// in order by one lane, so the running mean mix(c, s, 1/(i+1)) and every
// RNG key are exactly those of the straightforward kernel.
#include <hip/hip_runtime.h>

#include "fpt_device.h"
#include "fpt_internal.h"

namespace fpt {

enum : int { PH_IDLE = 0, PH_MARCH = 1, PH_EVENT = 2 };

struct PersistentCounters {
    int primary_steps, secondary_steps, shadow_steps, normal_evals, bounces;
};

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

// Wave-level fetch of `want` pixels: one atomic per wave, consecutive indices
// to the requesting lanes in lane order.  Must be reached by every lane.
__device__ __forceinline__ uint32_t fetch_pixel(bool want, uint32_t *counter) {
    const unsigned long long m = __ballot(want);
    uint32_t base = 0;
    if (m) {
        const int leader = __ffsll((long long)m) - 1;
        if (int(lane_id()) == leader) base = atomicAdd(counter, uint32_t(__popcll(m)));
        base = __shfl(base, leader, 64);
    }
    const unsigned long long below = m & ((1ull << lane_id()) - 1ull);
    return base + uint32_t(__popcll(below));
}

// Pixel index -> (x, y): 8x8 tiles, tiles row-major.
__device__ __forceinline__ void pixel_of(uint32_t k, uint32_t w, uint32_t &x, uint32_t &y) {
    const uint32_t tiles_x = (w + 7u) >> 3;
    const uint32_t tile = k >> 6, in = k & 63u;
    x = (tile % tiles_x) * 8u + (in & 7u);
    y = (tile / tiles_x) * 8u + (in >> 3);
}

template <bool PROFILE>
__global__ __launch_bounds__(256) void sdf_persistent_kernel(const FptRenderConfig *__restrict__ cfgp,
                                                             float4 *__restrict__ accum, uint32_t start,
                                                             uint32_t count, uint32_t *__restrict__ counter,
                                                             uint32_t slots, int event_batch, DeviceProfile *prof) {
    const FptRenderConfig &cfg = *cfgp;
    const uint32_t W = cfg.width, H = cfg.height;
    const uint32_t total = cfg.samples < 1u ? 1u : (cfg.samples > 512u ? 512u : cfg.samples);
    const uint32_t s_end = start + count < total ? start + count : total;
    const f3 cam = ld3(cfg.camera_position);
    const float focal = 1.0f / tanf(cfg.camera_fov / 2.0f * kPi / 180.0f);
    const float aa = 0.3f / fmaxf(float(W), float(H));
    const float maxd = cfg.render[4];
    const float maxd2 = maxd * maxd;
    const float far = maxd * 0.99f;
    const float far2 = far * far;
    const float min_dist = cfg.render[3];
    const float adapt = cfg.render[5];
    const bool check_t = cfg.sdf_id == FPT_SDF_PROGRAM && cfg.program_material[5] > 0.0f;
    int bounces = int(cfg.render[0]) < 8 ? int(cfg.render[0]) : 8;
    if (cfg.sdf_bounce_cap > 0u && int(cfg.sdf_bounce_cap) < bounces) bounces = int(cfg.sdf_bounce_cap);
    const int ni_first = int(cfg.render[1]);
    const int ni_next = int(cfg.render[1] / (cfg.render[5] * 2.0f + 1.0f));
    const bool sun_on = cfg.sun[0] == 1.0f;
    const f3 sun_dir0 = rotate_camera(mk(0.f, 0.f, 1.f), cfg.sun[1] * kPi / 180.0f, cfg.sun[2] * kPi / 180.0f);

    Work w = {0, 0};
    PersistentCounters pc = {0, 0, 0, 0, 0};
    int pixels_done = 0;

    // Lane state.
    int phase = PH_IDLE;
    uint32_t pix = 0, s = 0;
    UV uv = {0.f, 0.f};
    f3 acc = mk(0.f, 0.f, 0.f);
    // path
    int bounce = 0, local_ni = ni_first;
    float side = 1.0f, sky_mask = 0.0f;
    f3 light = mk(0.f, 0.f, 0.f), color = mk(1.f, 1.f, 1.f), gcol = mk(0.f, 0.f, 0.f), dr = mk(0.f, 0.f, 0.f);
    f3 hit_p = mk(0.f, 0.f, 0.f), hit_n = mk(0.f, 0.f, 0.f);
    Material mat;
    mat.rgb = mk(1.f, 1.f, 1.f);
    mat.roughness = 1.f; mat.specular = 0.f; mat.translucency = 0.f; mat.ior = 1.5f; mat.emission = 0.f;
    // march
    f3 p = mk(0.f, 0.f, 0.f), mdir = mk(0.f, 0.f, 1.f), mstart = mk(0.f, 0.f, 0.f);
    int it = 0, max_iter = 0;
    bool shadow = false;

    // Start the camera path of sample `s` of the current pixel.
    auto start_path = [&]() {
        const float frame = float(s);
        float jx, jy;
        random_point(aa, uv, frame, jx, jy);
        dr = rotate_camera_roll(normalize(mk(uv.x + jx, uv.y + jy, focal)), cfg.camera_yaw_pitch[0],
                                cfg.camera_yaw_pitch[1], cfg.camera_roll);
        const float gt = clampf((dr.y + 1.0f) / 2.0f, 0.f, 1.f);
        gcol = vmix(ld3(&cfg.background_gradient[3]), ld3(&cfg.background_gradient[0]), gt);
        bounce = 0;
        local_ni = ni_first;
        side = 1.0f;
        sky_mask = 0.0f;
        light = mk(0.f, 0.f, 0.f);
        color = mk(1.f, 1.f, 1.f);
        p = cam;
        mstart = cam;
        mdir = dr;
        it = 0;
        max_iter = local_ni < 360 ? local_ni : 360;
        shadow = false;
        phase = PH_MARCH;
    };

    // Take a pixel (or go idle).  Wave-uniform call.
    auto take_pixel = [&](bool want) {
        const uint32_t k = fetch_pixel(want, counter);
        if (!want) return;
        if (k >= slots) {
            phase = PH_IDLE;
            return;
        }
        uint32_t x, y;
        pixel_of(k, W, x, y);
        if (x >= W || y >= H) {  // tile padding: nothing to render, take another
            phase = PH_EVENT;     // handled as "needs a pixel" below
            pix = 0xffffffffu;
            return;
        }
        pix = y * W + x;
        uv = screen_uv(x, y, W, H);
        s = start;
        if (start != 0u) {
            float4 prev = accum[pix];
            acc = mk(prev.x, prev.y, prev.z);
        } else {
            acc = mk(0.f, 0.f, 0.f);
        }
        start_path();
    };

    take_pixel(true);
    // Tile-padding lanes ask again until every lane has a real pixel or none.
    while (true) {
        bool need = phase == PH_EVENT && pix == 0xffffffffu;
        if (!__any(need)) break;
        take_pixel(need);
    }

    while (true) {
        if (phase == PH_MARCH) {
            const float d = map_sdf(p, cfg, w);
            ++it;
            bool stop = !isfinite(d);
            if (!stop) {
                const float o = fabsf(d) * 0.99f;
                p = p + mdir * o;
                float lod = min_dist;
                {
                    const f3 df = mstart - p;  // lod_falloff 0.0002 (SM:3771-3776)
                    lod = mixf(min_dist, mixf(lod, 0.1f, dot(df, df) * 0.0002f), adapt);
                }
                if (check_t && user_material(p, cfg, w).translucency > 0.0f) lod = 0.0001f;
                stop = o < lod || maxd < o || it >= max_iter;
            }
            if (stop) phase = PH_EVENT;
        }
        const unsigned long long ev = __ballot(phase == PH_EVENT);
        const unsigned long long mar = __ballot(phase == PH_MARCH);
        if (!ev && !mar) break;
        if (!(__popcll(ev) >= event_batch || (ev && !mar))) continue;

        bool need_pixel = false;
        if (phase == PH_EVENT) {
            const float frame = float(s);
            bool path_done = false;
            bool do_scatter = false;
            if (shadow) {
                // End of the sun shadow march (SM:5856-5860).
                pc.shadow_steps += it;
                if (length(hit_p - p) > maxd * 0.99f) {
                    const float k = cfg.sun[3] * fmaxf(dot(hit_n, mdir), 0.0f) * mat.roughness * (1.0f - mat.translucency);
                    light = light + ld3(cfg.sun_color) * k;
                }
                do_scatter = true;
            } else {
                // End of a camera / bounce march (SM:5971-5989).
                if (bounce == 0) pc.primary_steps += it; else pc.secondary_steps += it;
                pc.bounces++;
                const f3 tv = p - cam;
                const float travel2 = dot(tv, tv);
                if (bounce == 0 && travel2 > maxd2) sky_mask = 1.0f;
                local_ni = ni_next;
                if (travel2 > far2) {
                    light = light + environment(dr, cfg);
                    path_done = true;
                } else {
                    hit_p = p;
                    mat = user_material(p, cfg, w);
                    if (mat.emission > 0.001f) light = light + mat.rgb * mat.emission;
                    hit_n = normal_at(p, cfg, w);
                    pc.normal_evals++;
                    if (sun_on) {
                        // Shadow ray (SM:5846-5856).
                        float h1 = hash13(uv.x, uv.y, madd_rn(frame, 5.0f, 1.0f));
                        float h2 = hash13(uv.x, uv.y, madd_rn(frame, 3.0f, 5.0f));
                        float sn, cs;
                        sincosf(h1 * 2.0f * kPi, &sn, &cs);
                        float m = sqrtf(h2) * cfg.sun[4];
                        mdir = rotate_camera(sun_dir0, cs * m, sn * m);
                        p = hit_p + hit_n * 0.001f;
                        mstart = p;
                        it = 0;
                        max_iter = ni_first < 360 ? ni_first : 360;
                        shadow = true;
                        phase = PH_MARCH;
                    } else {
                        do_scatter = true;
                    }
                }
            }
            if (do_scatter) {
                // Scatter (SM:5990-6021).
                shadow = false;
                f3 n = hit_n;
                const float fi = float(bounce);
                float r1 = hash13(uv.x, uv.y, madd_rn(frame, 1.37f, fi));
                float r2 = hash13(uv.x, uv.y, madd_rn(frame, 7.91f, fi));
                if (r1 > mat.translucency) {
                    f3 metal = reflect(dr, n);
                    f3 diffuse = random_vector(n, uv, madd_rn(frame, 13.37f, fi));
                    float q = (mat.ior - 1.0f) / (mat.ior + 1.0f);
                    float f0 = q * q;
                    float om = 1.0f - clampf(dot(n, -dr), 0.f, 1.f);
                    float om2 = om * om;
                    float fres = f0 + (1.0f - f0) * (om2 * om2 * om);
                    dr = normalize(vmix(metal, diffuse, mat.roughness));
                    if (r2 < fres * mat.specular) dr = metal;
                    else color = color * mat.rgb;
                } else {
                    float eta = side == 1.0f ? 1.0f / mat.ior : mat.ior;
                    if (side != 1.0f) n = -n;
                    float q = (mat.ior - 1.0f) / (mat.ior + 1.0f);
                    float f0 = q * q;
                    float om = 1.0f - clampf(dot(n, -dr), 0.f, 1.f);
                    float om2 = om * om;
                    float fres = f0 + (1.0f - f0) * (om2 * om2 * om);
                    f3 refr = refract(dr, n, eta);
                    f3 refl = reflect(dr, n);
                    if (dot(refr, refr) < 0.000001f || !isfinite(refr.x) || r2 < fres) {
                        dr = normalize(refl);
                    } else {
                        dr = normalize(refr);
                        side = -side;
                        color = color * vmix(mk(1.f, 1.f, 1.f), mat.rgb, 0.35f) * (1.0f - fres * 0.5f);
                    }
                }
                f3 rp = hit_p + n * (0.001f * signf(dot(dr, n)));
                bool stop = false;
                if (cfg.sdf_russian_roulette != 0u && bounce + 1 < bounces && float(bounce + 1) >= cfg.sdf_rr_start) {
                    float surv = clampf(fmaxf(color.x, fmaxf(color.y, color.z)), clampf(cfg.sdf_rr_min_prob, 0.01f, 1.0f),
                                        1.0f);
                    float rr = hash13(uv.x, uv.y, madd_rn(frame, 19.19f, __fmul_rn(fi, 3.17f)));
                    if (rr > surv) stop = true;
                    else color = color / surv;
                }
                ++bounce;
                if (stop || bounce >= bounces) {
                    path_done = true;
                } else {
                    p = rp;
                    mstart = rp;
                    mdir = dr;
                    it = 0;
                    max_iter = local_ni < 360 ? local_ni : 360;
                    phase = PH_MARCH;
                }
            }
            if (path_done) {
                f3 col = light * color;
                if (cfg.world[6] == 1.0f) col = vmix(col, gcol, sky_mask);
                col = vmin(col, 8.0f);
                acc = vmix(acc, col, 1.0f / float(s + 1u));
                ++s;
                if (s < s_end) {
                    start_path();
                } else {
                    accum[pix] = make_float4(acc.x, acc.y, acc.z, 1.0f);
                    ++pixels_done;
                    need_pixel = true;
                }
            }
        }
        // Wave-uniform pixel hand-out for lanes whose pixel is finished.
        take_pixel(need_pixel);
        while (true) {
            bool need = phase == PH_EVENT && pix == 0xffffffffu;
            if (!__any(need)) break;
            take_pixel(need);
        }
    }

    if (PROFILE) {
        auto wadd = [&](unsigned long long *dst, unsigned long long v) {
            for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
            if (lane_id() == 0u && v) atomicAdd(dst, v);
        };
        wadd(&prof->distance_evals, (unsigned long long)w.distance_evals);
        wadd(&prof->orbit_iterations, (unsigned long long)w.orbit_iterations);
        wadd(&prof->primary_steps, (unsigned long long)pc.primary_steps);
        wadd(&prof->secondary_steps, (unsigned long long)pc.secondary_steps);
        wadd(&prof->shadow_steps, (unsigned long long)pc.shadow_steps);
        wadd(&prof->normal_evals, (unsigned long long)pc.normal_evals);
        wadd(&prof->bounces, (unsigned long long)pc.bounces);
        wadd(&prof->pixels, (unsigned long long)pixels_done);
    }
}

hipError_t launch_sdf_persistent(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg, float4 *accum,
                                 uint32_t start, uint32_t count, uint32_t *counter, uint32_t workgroups,
                                 int event_batch, DeviceProfile *prof, hipStream_t stream) {
    // Pixel slots in 8x8 tile order (tiles padded to whole 8x8 blocks).
    const uint32_t tiles = ((hcfg.width + 7u) / 8u) * ((hcfg.height + 7u) / 8u);
    const uint32_t slots = tiles * 64u;
    hipError_t e = hipMemsetAsync(counter, 0, sizeof(uint32_t), stream);
    if (e != hipSuccess) return e;
    if (prof)
        hipLaunchKernelGGL(sdf_persistent_kernel<true>, dim3(workgroups), dim3(256), 0, stream, dcfg, accum, start,
                           count, counter, slots, event_batch, prof);
    else
        hipLaunchKernelGGL(sdf_persistent_kernel<false>, dim3(workgroups), dim3(256), 0, stream, dcfg, accum, start,
                           count, counter, slots, event_batch, nullptr);
    return hipGetLastError();
}

}  // namespace fpt
