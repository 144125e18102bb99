// SDF path-tracing kernels for gfx950 (MI355X).
//
// sdf_accumulate_kernel is the fused per-pixel path: camera -> march ->
// material -> analytic normal -> sun shadow march -> scatter -> next bounce
// -> running mean (Shaders.metal:5926-6026 inside 7757-7773).  Launch shape:
// 256-thread workgroups covering a 16x16 pixel tile, each 64-lane wave an
// 8x8 sub-tile so that the rays of one wave are spatially coherent and
// their march loops diverge as little as possible.
#include <hip/hip_runtime.h>

#include "fpt_device.h"
#include "fpt_internal.h"

namespace fpt {

// sunContributionWithSurface (SM:5846-5861).
__device__ f3 sun_contribution(f3 rp, UV xy, float seed, const Material &mat, f3 n, const FptRenderConfig &cfg,
                               Work &w, int &shadow_steps) {
    const float p1 = mat.roughness, p2 = 1.0f - mat.translucency;
    f3 ld = rotate_camera(mk(0.f, 0.f, 1.f), cfg.sun[1] * kPi / 180.0f, cfg.sun[2] * kPi / 180.0f);
    const f3 rp0 = rp;
    rp = rp + n * 0.001f;
    float h1 = hash13(xy.x, xy.y, madd_rn(seed, 5.0f, 1.0f));
    float h2 = hash13(xy.x, xy.y, madd_rn(seed, 3.0f, 5.0f));
    float s, c;
    sincosf(h1 * 2.0f * kPi, &s, &c);
    float m = sqrtf(h2) * cfg.sun[4];
    ld = rotate_camera(ld, c * m, s * m);
    int steps = 0;
    rp = march(ld, rp, int(cfg.render[1]), cfg.render[3], 0.0002f, cfg, w, steps);
    shadow_steps += steps;
    if (length(rp0 - rp) > cfg.render[4] * 0.99f) {
        float k = cfg.sun[3] * fmaxf(dot(n, ld), 0.0f) * p1 * p2;
        return ld3(cfg.sun_color) * k;
    }
    return mk(0.f, 0.f, 0.f);
}

struct PathCounts {
    int primary_steps, secondary_steps, shadow_steps, normal_evals, bounces;
};

// renderPath (SM:5926-6026); depth of field is not part of this library.
__device__ f3 render_path(UV xy, uint32_t sample_idx, const FptRenderConfig &cfg, Work &w, PathCounts &pc) {
    const float frame = float(sample_idx);
    const f3 cam = ld3(cfg.camera_position);
    f3 rp = cam;
    const float aa = 0.3f / fmaxf(float(cfg.width), float(cfg.height));
    float jx, jy;
    random_point(aa, xy, frame, jx, jy);
    const float focal = 1.0f / tanf(cfg.camera_fov / 2.0f * kPi / 180.0f);
    f3 dr = rotate_camera_roll(normalize(mk(xy.x + jx, xy.y + jy, focal)), cfg.camera_yaw_pitch[0],
                               cfg.camera_yaw_pitch[1], cfg.camera_roll);
    int local_ni = int(cfg.render[1]);
    float side = 1.0f;
    f3 light = mk(0.f, 0.f, 0.f), color = mk(1.f, 1.f, 1.f);
    float sky_mask = 0.0f;
    const float gt = clampf((dr.y + 1.0f) / 2.0f, 0.f, 1.f);
    const f3 gcol = vmix(ld3(&cfg.background_gradient[3]), ld3(&cfg.background_gradient[0]), gt);
    int bounces = int(cfg.render[0]) < 8 ? int(cfg.render[0]) : 8;
    if (cfg.sdf_bounce_cap > 0u && int(cfg.sdf_bounce_cap) < bounces) bounces = int(cfg.sdf_bounce_cap);
    const float maxd = cfg.render[4];
    const float maxd2 = maxd * maxd;
    const float far = maxd * 0.99f;
    const float far2 = far * far;
    for (int i = 0; i < bounces; ++i) {
        int steps = 0;
        rp = march(dr, rp, local_ni, cfg.render[3], 0.0002f, cfg, w, steps);
        if (i == 0) pc.primary_steps += steps; else pc.secondary_steps += steps;
        pc.bounces++;
        f3 tv = rp - cam;
        float travel2 = dot(tv, tv);
        if (i == 0 && travel2 > maxd2) sky_mask = 1.0f;
        local_ni = int(cfg.render[1] / (cfg.render[5] * 2.0f + 1.0f));
        if (travel2 > far2) {
            light = light + environment(dr, cfg);
            break;
        }
        Material mat = user_material(rp, cfg, w);
        if (mat.emission > 0.001f) light = light + mat.rgb * mat.emission;
        f3 n = normal_at(rp, cfg, w);
        pc.normal_evals++;
        if (cfg.sun[0] == 1.0f) light = light + sun_contribution(rp, xy, frame, mat, n, cfg, w, pc.shadow_steps);
        const float fi = float(i);
        float r1 = hash13(xy.x, xy.y, madd_rn(frame, 1.37f, fi));
        float r2 = hash13(xy.x, xy.y, madd_rn(frame, 7.91f, fi));
        if (r1 > mat.translucency) {
            f3 metal = reflect(dr, n);
            f3 diffuse = random_vector(n, xy, madd_rn(frame, 13.37f, fi));
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
        rp = rp + n * (0.001f * signf(dot(dr, n)));
        if (cfg.sdf_russian_roulette != 0u && i + 1 < bounces && float(i + 1) >= cfg.sdf_rr_start) {
            float surv = clampf(fmaxf(color.x, fmaxf(color.y, color.z)), clampf(cfg.sdf_rr_min_prob, 0.01f, 1.0f), 1.0f);
            float rr = hash13(xy.x, xy.y, madd_rn(frame, 19.19f, __fmul_rn(fi, 3.17f)));
            if (rr > surv) break;
            color = color / surv;
        }
    }
    f3 col = light * color;
    if (cfg.world[6] == 1.0f) col = vmix(col, gcol, sky_mask);
    return vmin(col, 8.0f);
}

// 16x16 pixel tile per 256-thread workgroup; each wave an 8x8 sub-tile.
__device__ __forceinline__ void tile_pixel(uint32_t &x, uint32_t &y) {
    const uint32_t t = threadIdx.x;
    const uint32_t wave = t >> 6, lane = t & 63u;
    x = blockIdx.x * 16u + (wave & 1u) * 8u + (lane & 7u);
    y = blockIdx.y * 16u + (wave >> 1) * 8u + (lane >> 3);
}

__device__ __forceinline__ void wave_add_u64(unsigned long long *dst, unsigned long long v) {
    // Reduce across the 64-lane wave, one atomic per wave.
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63u) == 0u && v) atomicAdd(dst, v);
}

template <bool PROFILE>
__global__ __launch_bounds__(256) void sdf_accumulate_kernel(const FptRenderConfig *__restrict__ cfgp,
                                                             float4 *__restrict__ accum, uint32_t start,
                                                             uint32_t count, DeviceProfile *prof) {
    const FptRenderConfig &cfg = *cfgp;
    uint32_t x, y;
    tile_pixel(x, y);
    const bool active = x < cfg.width && y < cfg.height;
    Work w = {0, 0};
    PathCounts pc = {0, 0, 0, 0, 0};
    if (active) {
        const UV uv = screen_uv(x, y, cfg.width, cfg.height);
        const size_t idx = size_t(y) * cfg.width + x;
        f3 c = mk(0.f, 0.f, 0.f);
        if (start != 0u) {
            float4 prev = accum[idx];
            c = mk(prev.x, prev.y, prev.z);
        }
        uint32_t total = cfg.samples < 1u ? 1u : (cfg.samples > 512u ? 512u : cfg.samples);
        uint32_t end = start + count < total ? start + count : total;
        for (uint32_t s = start; s < end; ++s) {
            f3 sc = render_path(uv, s, cfg, w, pc);
            c = vmix(c, sc, 1.0f / float(s + 1u));
        }
        accum[idx] = make_float4(c.x, c.y, c.z, 1.0f);
    }
    if (PROFILE) {
        wave_add_u64(&prof->distance_evals, (unsigned long long)w.distance_evals);
        wave_add_u64(&prof->orbit_iterations, (unsigned long long)w.orbit_iterations);
        wave_add_u64(&prof->primary_steps, (unsigned long long)pc.primary_steps);
        wave_add_u64(&prof->secondary_steps, (unsigned long long)pc.secondary_steps);
        wave_add_u64(&prof->shadow_steps, (unsigned long long)pc.shadow_steps);
        wave_add_u64(&prof->normal_evals, (unsigned long long)pc.normal_evals);
        wave_add_u64(&prof->bounces, (unsigned long long)pc.bounces);
        wave_add_u64(&prof->pixels, active ? 1ull : 0ull);
    }
}

// Un-jittered primary ray + normal per pixel (per-ray parity harness).
__global__ __launch_bounds__(256) void trace_primary_kernel(const FptRenderConfig *__restrict__ cfgp,
                                                            FptRayHit *__restrict__ hits) {
    const FptRenderConfig &cfg = *cfgp;
    uint32_t x, y;
    tile_pixel(x, y);
    if (x >= cfg.width || y >= cfg.height) return;
    const UV uv = screen_uv(x, y, cfg.width, cfg.height);
    const f3 cam = ld3(cfg.camera_position);
    const float focal = 1.0f / tanf(cfg.camera_fov / 2.0f * kPi / 180.0f);
    f3 dr = rotate_camera_roll(normalize(mk(uv.x, uv.y, focal)), cfg.camera_yaw_pitch[0], cfg.camera_yaw_pitch[1],
                               cfg.camera_roll);
    Work w = {0, 0};
    int steps = 0;
    f3 rp = march(dr, cam, int(cfg.render[1]), cfg.render[3], 0.0002f, cfg, w, steps);
    f3 tv = rp - cam;
    float travel2 = dot(tv, tv);
    float far = cfg.render[4] * 0.99f;
    bool hit = travel2 <= far * far;
    f3 n = hit ? normal_at(rp, cfg, w) : mk(0.f, 0.f, 0.f);
    FptRayHit h;
    h.t = sqrtf(travel2);
    h.position[0] = rp.x; h.position[1] = rp.y; h.position[2] = rp.z;
    h.normal[0] = n.x; h.normal[1] = n.y; h.normal[2] = n.z;
    h.hit = hit ? 1u : 0u;
    h.steps = uint32_t(steps);
    h._pad0 = 0u;
    hits[size_t(y) * cfg.width + x] = h;
}

// Distance field samples: {distance, radius, derivative, iterations} of the
// first fractal leaf of the program (or the scene distance alone).
__global__ __launch_bounds__(256) void field_sample_kernel(const FptRenderConfig *__restrict__ cfgp,
                                                           const float4 *__restrict__ pts, size_t n,
                                                           FptMandelbulberFieldSample *__restrict__ out) {
    const FptRenderConfig &cfg = *cfgp;
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 q = pts[i];
    f3 p = mk(q.x, q.y, q.z);
    Work w = {0, 0};
    FptMandelbulberFieldSample s;
    s.distance = map_sdf(p, cfg, w);
    s.radius = 0.f;
    s.derivative = 0.f;
    s.iterations = 0.f;
    if (is_program(cfg)) {
        // Replay transforms up to the first fractal leaf for its orbit stats.
        float ds = 1.0f;
        const uint32_t cnt = prog_count(cfg);
        for (uint32_t k = 0; k < cnt; ++k) {
            const FptSdfInstruction &op = cfg.sdf_program[k];
            if (apply_transform(op, p, ds)) continue;
            if (is_leaf(op.opcode)) {
                Leaf<float> lf = eval_leaf<float>(op.opcode, v3<float>{p.x, p.y, p.z}, op.data, nullptr);
                s.radius = lf.radius;
                s.derivative = lf.derivative;
                s.iterations = float(lf.iterations);
                break;
            }
        }
    }
    out[i] = s;
}

// Program surface (distance + analytic gradient) at points (parity hook).
__global__ __launch_bounds__(256) void surface_sample_kernel(const FptRenderConfig *__restrict__ cfgp,
                                                             const float4 *__restrict__ pts, size_t n,
                                                             float4 *__restrict__ out) {
    const FptRenderConfig &cfg = *cfgp;
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 q = pts[i];
    f3 g;
    float d = program_surface(cfg.sdf_program, prog_count(cfg), mk(q.x, q.y, q.z), g);
    out[i] = make_float4(d, g.x, g.y, g.z);
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
static dim3 tile_grid(uint32_t w, uint32_t h) { return dim3((w + 15u) / 16u, (h + 15u) / 16u, 1u); }

hipError_t launch_sdf_accumulate(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg, float4 *accum,
                                 uint32_t start, uint32_t count, DeviceProfile *prof, hipStream_t stream) {
    dim3 grid = tile_grid(hcfg.width, hcfg.height);
    if (prof)
        hipLaunchKernelGGL(sdf_accumulate_kernel<true>, grid, dim3(256), 0, stream, dcfg, accum, start, count, prof);
    else
        hipLaunchKernelGGL(sdf_accumulate_kernel<false>, grid, dim3(256), 0, stream, dcfg, accum, start, count,
                           nullptr);
    return hipGetLastError();
}

hipError_t launch_trace_primary(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg, FptRayHit *hits,
                                hipStream_t stream) {
    hipLaunchKernelGGL(trace_primary_kernel, tile_grid(hcfg.width, hcfg.height), dim3(256), 0, stream, dcfg, hits);
    return hipGetLastError();
}

hipError_t launch_field_sample(const FptRenderConfig *dcfg, const float4 *pts, size_t n,
                               FptMandelbulberFieldSample *out, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(field_sample_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream, dcfg, pts, n, out);
    return hipGetLastError();
}

hipError_t launch_surface_sample(const FptRenderConfig *dcfg, const float4 *pts, size_t n, float4 *out,
                                 hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(surface_sample_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream, dcfg, pts, n,
                       out);
    return hipGetLastError();
}

}  // namespace fpt
