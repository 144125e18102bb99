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
#include "post.h"

namespace fpt {

// sunContributionWithSurface (SM:5846-5861); K: the scene kinds this
// instance serves (SCENE_ANY decides at run time).
template <int K>
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
    rp = march_any<K>(ld, rp, int(cfg.render[1]), cfg.render[3], 0.0002f, cfg, w, steps);
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

// renderPath (SM:5926-6026); K as for sun_contribution.
template <int K>
__device__ f3 render_path(UV xy, uint32_t sample_idx, const FptRenderConfig &cfg, Work &w, PathCounts &pc) {
    const float frame = float(sample_idx);
    const f3 cam = ld3(cfg.camera_position);
    f3 rp = cam;
    const float aa = 0.3f / fmaxf(float(cfg.width), float(cfg.height));
    float jx, jy;
    random_point(aa, xy, frame, jx, jy);
    f3 dr;
    if (!camera_ray_any<K>(xy.x + jx, xy.y + jy, cfg, dr)) return mk(0.f, 0.f, 0.f);
    apply_dof(xy, frame, cfg, rp, dr);
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
        rp = march_any<K>(dr, rp, local_ni, cfg.render[3], 0.0002f, cfg, w, steps);
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
        Material mat = material_any<K>(rp, cfg, w);
        if (mat.emission > 0.001f) light = light + mat.rgb * mat.emission;
        f3 n = normal_any<K>(rp, cfg, w);
        pc.normal_evals++;
        if (cfg.sun[0] == 1.0f) light = light + sun_contribution<K>(rp, xy, frame, mat, n, cfg, w, pc.shadow_steps);
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
                color = color * (vmix(mk(1.f, 1.f, 1.f), mat.rgb, 0.35f) * (1.0f - fres * 0.5f));
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

// Minimum resident waves per SIMD for the imported-Mandelbulber instance.
// Its KIFS orbit reads the fold planes from the config with scalar loads,
// so more resident waves hide them: measured on ifs-20.fract at 1080p,
// 8 spp: 3 waves (134 VGPRs, no hint) 838 ms, 4 waves 691 ms, 5 waves (96
// VGPRs, 128 B spilled per lane) 616 ms, identical images.
#ifndef FPT_MANDELBULBER_WAVES
#define FPT_MANDELBULBER_WAVES 5
#endif

// K: SCENE_SDF (typed programs and built-in fractals) or SCENE_MB
// (imported Mandelbulber scenes), chosen per config at launch.
template <bool PROFILE, int K>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(K == SCENE_MB ? FPT_MANDELBULBER_WAVES : 1)))
void sdf_accumulate_kernel(const FptRenderConfig *__restrict__ cfgp,
                                                             float4 *__restrict__ accum, uint32_t start,
                                                             uint32_t count, DeviceProfile *prof) {
    const FptRenderConfig &cfg = *cfgp;
    uint32_t x, y;
    tile_pixel(x, y);
    const bool active = x < cfg.width && y < cfg.height;
    Work w = {0, 0};
    PathCounts pc = {0, 0, 0, 0, 0};
    if (active) {
        const UV uv = screen_uv_any<K>(x, y, cfg);
        const size_t idx = size_t(y) * cfg.width + x;
        f3 c = mk(0.f, 0.f, 0.f);
        if (start != 0u) {
            float4 prev = accum[idx];
            c = mk(prev.x, prev.y, prev.z);
        }
        uint32_t total = cfg.samples < 1u ? 1u : (cfg.samples > 512u ? 512u : cfg.samples);
        uint32_t end = start + count < total ? start + count : total;
        for (uint32_t s = start; s < end; ++s) {
            f3 sc = render_path<K>(uv, s, cfg, w, pc);
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

// Work of one path sample per pixel (sample 0 of the frame's sequence, the
// whole path: camera, bounce and sun marches and the per-hit work), as
// distance evaluations + orbit iterations, into cost[y * width + x].  The
// persistent kernel hands out tiles in the order of these costs.
template <int K>
__global__ __launch_bounds__(256) void pixel_cost_kernel(const FptRenderConfig *__restrict__ cfgp,
                                                         float *__restrict__ cost) {
    const FptRenderConfig &cfg = *cfgp;
    uint32_t x, y;
    tile_pixel(x, y);
    if (x >= cfg.width || y >= cfg.height) return;
    Work w = {0, 0};
    PathCounts pc = {0, 0, 0, 0, 0};
    (void)render_path<K>(screen_uv_any<K>(x, y, cfg), 0u, cfg, w, pc);
    cost[size_t(y) * cfg.width + x] = float(w.distance_evals) + float(w.orbit_iterations);
}

// Un-jittered primary ray + normal per pixel (per-ray parity harness).
__global__ __launch_bounds__(256) void trace_primary_kernel(const FptRenderConfig *__restrict__ cfgp,
                                                            FptRayHit *__restrict__ hits) {
    const FptRenderConfig &cfg = *cfgp;
    uint32_t x, y;
    tile_pixel(x, y);
    if (x >= cfg.width || y >= cfg.height) return;
    const UV uv = screen_uv_any(x, y, cfg);
    const f3 cam = ld3(cfg.camera_position);
    f3 dr = mk(0.f, 0.f, 0.f);
    const bool seen = camera_ray_any(uv.x, uv.y, cfg, dr);
    Work w = {0, 0};
    int steps = 0;
    f3 rp = seen ? march_any(dr, cam, int(cfg.render[1]), cfg.render[3], 0.0002f, cfg, w, steps) : cam;
    f3 tv = rp - cam;
    float travel2 = dot(tv, tv);
    float far = cfg.render[4] * 0.99f;
    bool hit = seen && travel2 <= far * far;
    f3 n = hit ? normal_any(rp, cfg, w) : mk(0.f, 0.f, 0.f);
    FptRayHit h;
    h.t = sqrtf(travel2);
    h.position[0] = rp.x; h.position[1] = rp.y; h.position[2] = rp.z;
    h.normal[0] = n.x; h.normal[1] = n.y; h.normal[2] = n.z;
    h.hit = hit ? 1u : 0u;
    h.steps = uint32_t(steps);
    h._pad0 = 0u;
    hits[size_t(y) * cfg.width + x] = h;
}

// Distance field samples {distance, radius, derivative, iterations}: the
// Mandelbulber native KIFS field for sdf_id MANDELBULBER, otherwise the
// scene distance with the orbit statistics of the program's first fractal
// leaf.
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
    if (cfg.sdf_id == FPT_SDF_MANDELBULBER) {  // mandelbulber_field_sample_kernel (SM:1001-1008)
        const float4 k = kifs_field_sample(p, cfg, 1);
        s.distance = k.x;
        s.radius = k.y;
        s.derivative = k.z;
        s.iterations = fabsf(k.w);
        out[i] = s;
        return;
    }
    s.distance = map_scene(p, cfg, w);
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

// typed_soa_validation_kernel (SM:5106-5154): the typed structure-of-arrays
// table against the interpreted program at hashed points of the voxel
// bounds (lowbias32, SM:4992-4999), distance within 2e-5 (1 + |d|) and
// gradient within 1e-4 per component.  Counts are reduced across the wave
// first: one atomic per wave and counter.
__device__ __forceinline__ uint32_t validation_hash(uint32_t v) {  // SM:4992-4999
    v ^= v >> 16u;
    v *= 0x7feb352du;
    v ^= v >> 15u;
    v *= 0x846ca68bu;
    v ^= v >> 16u;
    return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, uint32_t(__shfl_xor(int(v), o, 64)));
    return v;
}
__global__ __launch_bounds__(256) void typed_soa_validation_kernel(const FptRenderConfig *__restrict__ cfgp,
                                                                   uint32_t n, uint32_t *__restrict__ counts) {
    const FptRenderConfig &cfg = *cfgp;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    bool dfail = false, gfail = false;
    uint32_t dbits = 0u, gbits = 0u;
    if (gid < n) {
        const uint32_t hx = validation_hash(gid * 3u + 0x9e3779b9u);
        const uint32_t hy = validation_hash(gid * 3u + 0x243f6a88u);
        const uint32_t hz = validation_hash(gid * 3u + 0xb7e15162u);
        const f3 unit = mk(float(hx), float(hy), float(hz)) * (1.0f / 4294967295.0f);
        const f3 lo = ld3(cfg.voxel_bounds_min), hi = ld3(cfg.voxel_bounds_max);
        f3 p = mk(mixf(lo.x, hi.x, unit.x), mixf(lo.y, hi.y, unit.y), mixf(lo.z, hi.z, unit.z));
        if (gid == 0u) p = mk(0.f, 0.f, 0.f);
        if (gid == 1u) p = lo;
        if (gid == 2u) p = hi;
        if (gid == 3u) p = (lo + hi) * 0.5f;
        const float ref = program_eval<false>(cfg.sdf_program, prog_count(cfg), p, nullptr, nullptr);
        const float direct = typed_soa_distance(p, cfg);
        const float derr = fabsf(ref - direct);
        dfail = !isfinite(ref) || !isfinite(direct) || derr > 2.0e-5f * (1.0f + fabsf(ref));
        dbits = __float_as_uint(derr);
        f3 gr, gd;
        program_surface(cfg.sdf_program, prog_count(cfg), p, gr);
        typed_soa_surface(p, cfg, gd);
        const float gerr = fmaxf(fabsf(gr.x - gd.x), fmaxf(fabsf(gr.y - gd.y), fabsf(gr.z - gd.z)));
        gfail = !finite3(gr) || !finite3(gd) || gerr > 1.0e-4f;
        gbits = __float_as_uint(gerr);
    }
    const uint32_t nd = uint32_t(__popcll(__ballot(dfail))), ng = uint32_t(__popcll(__ballot(gfail)));
    dbits = wave_max_u32(dbits);
    gbits = wave_max_u32(gbits);
    if ((threadIdx.x & 63u) == 0u) {
        if (nd) atomicAdd(&counts[0], nd);
        if (ng) atomicAdd(&counts[1], ng);
        atomicMax(&counts[2], dbits);
        atomicMax(&counts[3], gbits);
    }
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
static dim3 tile_grid(uint32_t w, uint32_t h) { return dim3((w + 15u) / 16u, (h + 15u) / 16u, 1u); }

// ---------------------------------------------------------------------------
// Diagnostic images (sdfDiagnostic, Shaders.metal:6827-6985, 7210-7373):
// one 8-bit RGBA value per pixel, flipped like the present kernel.
// ---------------------------------------------------------------------------
__device__ __forceinline__ f3 diag_scalar(float v, float scale) {  // SM:6842-6845
    const float t = clampf(v / fmaxf(scale, 1.0f), 0.f, 1.f);
    return mk(t, t, t);
}
__device__ __forceinline__ f3 diag_hdr(f3 c, const FptRenderConfig &cfg) {  // SM:6838-6840
    return post_process(vmin(vmax0(c), 8.0f), cfg);
}
__device__ __forceinline__ f3 diag_normal(f3 n) {  // SM:6834-6836
    return mk(clampf(n.x * 0.5f + 0.5f, 0.f, 1.f), clampf(n.y * 0.5f + 0.5f, 0.f, 1.f),
              clampf(n.z * 0.5f + 0.5f, 0.f, 1.f));
}
__device__ __forceinline__ f3 diag_depth(float travel, bool hit, const FptRenderConfig &cfg,
                                         const FptDiagnosticConfig &dg) {  // SM:6827-6832
    if (!hit) return mk(0.f, 0.f, 0.f);
    const float md = dg.max_distance > 0.001f ? dg.max_distance : cfg.render[4];
    const float v = 1.0f - clampf(travel / fmaxf(md, 0.001f), 0.f, 1.f);
    return mk(v, v, v);
}

__device__ __forceinline__ f3 jittered_camera_ray(UV xy, float frame, const FptRenderConfig &cfg) {
    const float focal = 1.0f / tanf(cfg.camera_fov / 2.0f * kPi / 180.0f);
    const float aa = 0.3f / fmaxf(float(cfg.width), float(cfg.height));
    float jx, jy;
    random_point(aa, xy, frame, jx, jy);
    return rotate_camera_roll(normalize(mk(xy.x + jx, xy.y + jy, focal)), cfg.camera_yaw_pitch[0],
                              cfg.camera_yaw_pitch[1], cfg.camera_roll);
}

__device__ __forceinline__ int bounce_limit(const FptRenderConfig &cfg) {
    int b = int(cfg.render[0]) < 8 ? int(cfg.render[0]) : 8;
    if (cfg.sdf_bounce_cap > 0u && int(cfg.sdf_bounce_cap) < b) b = int(cfg.sdf_bounce_cap);
    return b;
}

// sdfPathCostDiagnostic (SM:6919-6985): mean march steps, shadow steps,
// normal evaluations or bounces per path over min(spp, 8) samples.
__device__ f3 path_cost_diagnostic(UV xy, const FptRenderConfig &cfg, const FptDiagnosticConfig &dg, Work &w) {
    const uint32_t samples = cfg.samples < 1u ? 1u : (cfg.samples > 8u ? 8u : cfg.samples);
    float t_primary = 0.f, t_shadow = 0.f, t_normal = 0.f, t_bounces = 0.f;
    const f3 cam = ld3(cfg.camera_position);
    const float far = cfg.render[4] * 0.99f, far2 = far * far;
    const int bounces = bounce_limit(cfg);
    for (uint32_t si = 0; si < samples; ++si) {
        const float frame = float(si);
        f3 rp = cam;
        f3 dr = jittered_camera_ray(xy, frame, cfg);
        apply_dof(xy, frame, cfg, rp, dr);  // SM:6933-6939, 7222-7228
        int local_ni = int(cfg.render[1]);
        for (int i = 0; i < bounces; ++i) {
            int steps = 0;
            rp = march_counted_any(dr, rp, local_ni, cfg.render[3], 0.0002f, cfg, w, steps);
            t_primary += float(steps);
            local_ni = int(cfg.render[1] / (cfg.render[5] * 2.0f + 1.0f));
            const f3 tv = rp - cam;
            if (dot(tv, tv) > far2) break;
            const Material mat = material_any(rp, cfg, w);
            t_bounces += 1.0f;
            t_normal += 1.0f;
            const f3 n = normal_any(rp, cfg, w);
            if (cfg.sun[0] == 1.0f) {  // sunStepCountWithNormal (SM:6910-6917)
                f3 ld = rotate_camera(mk(0.f, 0.f, 1.f), cfg.sun[1] * kPi / 180.0f, cfg.sun[2] * kPi / 180.0f);
                const float h1 = hash13(xy.x, xy.y, madd_rn(frame, 5.0f, 1.0f));
                const float h2 = hash13(xy.x, xy.y, madd_rn(frame, 3.0f, 5.0f));
                float sn, cs;
                sincosf(h1 * 2.0f * kPi, &sn, &cs);
                const float m = sqrtf(h2) * cfg.sun[4];
                ld = rotate_camera(ld, cs * m, sn * m);
                int sst = 0;
                (void)march_counted_any(ld, rp + n * 0.001f, int(cfg.render[1]), cfg.render[3], 0.0002f, cfg, w,
                                        sst);
                t_shadow += float(sst);
            }
            const float fi = float(i);
            const float r1 = hash13(xy.x, xy.y, madd_rn(frame, 1.37f, fi));
            const float r2 = hash13(xy.x, xy.y, madd_rn(frame, 7.91f, fi));
            const float q = (mat.ior - 1.0f) / (mat.ior + 1.0f);
            const float f0 = q * q;
            const float om = 1.0f - clampf(dot(n, -dr), 0.f, 1.f);
            const float om2 = om * om;
            const float fres = f0 + (1.0f - f0) * (om2 * om2 * om);
            if (r1 > mat.translucency) {
                const f3 metal = reflect(dr, n);
                const f3 diffuse = random_vector(n, xy, madd_rn(frame, 13.37f, fi));
                dr = normalize(vmix(metal, diffuse, mat.roughness));
                if (r2 < fres * mat.specular) dr = metal;
            } else {
                const f3 refr = refract(dr, n, 1.0f / mat.ior);
                const f3 refl = reflect(dr, n);
                dr = (dot(refr, refr) < 0.000001f || !isfinite(refr.x) || r2 < fres) ? normalize(refl)
                                                                                    : normalize(refr);
            }
            rp = rp + n * (0.001f * signf(dot(dr, n)));
        }
    }
    const float inv = 1.0f / float(samples);
    const float steps_scale = fmaxf(cfg.render[1] * fmaxf(cfg.render[0], 1.0f), 1.0f);
    if (dg.mode == FPT_DIAGNOSTIC_SDF_PRIMARY_STEPS) return diag_scalar(t_primary * inv, steps_scale);
    if (dg.mode == FPT_DIAGNOSTIC_SDF_SHADOW_STEPS) return diag_scalar(t_shadow * inv, steps_scale);
    if (dg.mode == FPT_DIAGNOSTIC_SDF_NORMAL_EVALS)
        return diag_scalar(t_normal * inv, fmaxf(cfg.render[0] * 2.0f, 1.0f));
    return diag_scalar(t_bounces * inv, fmaxf(cfg.render[0], 1.0f));
}

// sdfBounceContributionDiagnostic (SM:7210-7279): emission + sun light of
// bounce `_pad0` only, times the throughput that reaches it.
__device__ f3 bounce_contribution_diagnostic(UV xy, const FptRenderConfig &cfg, const FptDiagnosticConfig &dg,
                                             Work &w) {
    const uint32_t target = dg._pad0;
    const uint32_t samples = cfg.samples < 1u ? 1u : (cfg.samples > 8u ? 8u : cfg.samples);
    const f3 cam = ld3(cfg.camera_position);
    const float far = cfg.render[4] * 0.99f, far2 = far * far;
    const int bounces = bounce_limit(cfg);
    f3 total = mk(0.f, 0.f, 0.f);
    int unused = 0;
    for (uint32_t si = 0; si < samples; ++si) {
        const float frame = float(si);
        f3 rp = cam;
        f3 dr = jittered_camera_ray(xy, frame, cfg);
        apply_dof(xy, frame, cfg, rp, dr);  // SM:6933-6939, 7222-7228
        int local_ni = int(cfg.render[1]);
        f3 thr = mk(1.f, 1.f, 1.f);
        for (int i = 0; i < bounces; ++i) {
            int steps = 0;
            rp = march_any(dr, rp, local_ni, cfg.render[3], 0.0002f, cfg, w, steps);
            const f3 tv = rp - cam;
            local_ni = int(cfg.render[1] / (cfg.render[5] * 2.0f + 1.0f));
            if (dot(tv, tv) > far2) {
                if (uint32_t(i) == target) total = total + environment(dr, cfg) * thr;
                break;
            }
            const Material mat = material_any(rp, cfg, w);
            const f3 n = normal_any(rp, cfg, w);
            f3 contrib = mk(0.f, 0.f, 0.f);
            if (mat.emission > 0.001f) contrib = contrib + mat.rgb * mat.emission;
            if (cfg.sun[0] == 1.0f) contrib = contrib + sun_contribution<SCENE_ANY>(rp, xy, frame, mat, n, cfg, w, unused);
            if (uint32_t(i) == target) total = total + contrib * thr;
            const float fi = float(i);
            const float r1 = hash13(xy.x, xy.y, madd_rn(frame, 1.37f, fi));
            const float r2 = hash13(xy.x, xy.y, madd_rn(frame, 7.91f, fi));
            const float q = (mat.ior - 1.0f) / (mat.ior + 1.0f);
            const float f0 = q * q;
            const float om = 1.0f - clampf(dot(n, -dr), 0.f, 1.f);
            const float om2 = om * om;
            const float fres = f0 + (1.0f - f0) * (om2 * om2 * om);
            if (r1 > mat.translucency) {
                const f3 metal = reflect(dr, n);
                const f3 diffuse = random_vector(n, xy, madd_rn(frame, 13.37f, fi));
                dr = normalize(vmix(metal, diffuse, mat.roughness));
                if (r2 >= fres * mat.specular) thr = thr * mat.rgb;
                else dr = metal;
            } else {
                const f3 refr = refract(dr, n, 1.0f / mat.ior);
                const f3 refl = reflect(dr, n);
                if (dot(refr, refr) < 0.000001f || !isfinite(refr.x) || r2 < fres) {
                    dr = normalize(refl);
                } else {
                    dr = normalize(refr);
                    thr = thr * (vmix(mk(1.f, 1.f, 1.f), mat.rgb, 0.35f) * (1.0f - fres * 0.5f));
                }
            }
            rp = rp + n * (0.001f * signf(dot(dr, n)));
        }
    }
    return diag_hdr(total / float(samples), cfg);
}

__device__ f3 sdf_diagnostic(UV xy, const FptRenderConfig &cfg, const FptDiagnosticConfig &dg, Work &w) {
    const uint32_t mode = dg.mode;
    if (mode == FPT_DIAGNOSTIC_SDF_BOUNCE_CONTRIBUTION) return bounce_contribution_diagnostic(xy, cfg, dg, w);
    if (mode == FPT_DIAGNOSTIC_SDF_PRIMARY_STEPS || mode == FPT_DIAGNOSTIC_SDF_SHADOW_STEPS ||
        mode == FPT_DIAGNOSTIC_SDF_NORMAL_EVALS || mode == FPT_DIAGNOSTIC_SDF_BOUNCES)
        return path_cost_diagnostic(xy, cfg, dg, w);
    if (mode == FPT_DIAGNOSTIC_PATH_FINAL) {
        const uint32_t samples = cfg.samples < 16u ? 16u : (cfg.samples > 64u ? 64u : cfg.samples);
        f3 c = mk(0.f, 0.f, 0.f);
        PathCounts pc = {0, 0, 0, 0, 0};
        for (uint32_t i = 0; i < samples; ++i) c = c + render_path<SCENE_ANY>(xy, i, cfg, w, pc);
        return post_process(c / float(samples), cfg);
    }
    // Un-jittered camera ray with the diagnostic's own march (SM:7304-7339);
    // Mandelbulber scenes use their camera and marchMandelbulber.
    const f3 ro = ld3(cfg.camera_position);
    f3 rd;
    if (!camera_ray_any(xy.x, xy.y, cfg, rd)) return mk(0.f, 0.f, 0.f);
    const int requested = int(fmaxf(cfg.render[1], 32.0f));
    f3 rp = ro;
    bool hit = false;
    if (is_mandelbulber(cfg)) {
        int unused = 0;
        rp = march_mandelbulber(rd, ro, requested < 10000 ? requested : 10000, cfg, w, unused, &hit);
    } else {
        const int steps = requested < 420 ? requested : 420;
        const float thr = fmaxf(cfg.render[3], 0.0005f);
        for (int i = 0; i < steps; ++i) {
            const float d = map_scene(rp, cfg, w);
            if (!isfinite(d)) break;
            if (d < thr && length(rp - ro) > 0.001f) {
                hit = true;
                break;
            }
            rp = rp + rd * (fabsf(d) * 0.99f);
            if (length(rp - ro) > cfg.render[4]) break;
        }
    }
    if (mode == FPT_DIAGNOSTIC_HIT_MASK) return hit ? mk(1.f, 1.f, 1.f) : mk(0.f, 0.f, 0.f);
    if (mode == FPT_DIAGNOSTIC_DEPTH) return diag_depth(length(rp - ro), hit, cfg, dg);
    if (!hit) return mk(0.f, 0.f, 0.f);
    if (mode == FPT_DIAGNOSTIC_DIFFUSE_NORMAL) {
        const f3 n = normal_any(rp, cfg, w);
        const f3 fwd = rotate_camera_roll(mk(0.f, 0.f, 1.f), cfg.camera_yaw_pitch[0], cfg.camera_yaw_pitch[1],
                                          cfg.camera_roll);
        const float k = 0.12f + 0.88f * fmaxf(dot(n, -fwd), 0.0f);
        return mk(k, k, k);
    }
    // Colour index / palette position need the reference's generated
    // Mandelbulber material code; without it the reference returns black.
    if (mode == FPT_DIAGNOSTIC_MANDEL_COLOR_INDEX || mode == FPT_DIAGNOSTIC_MANDEL_PALETTE_POSITION)
        return mk(0.f, 0.f, 0.f);
    const Material mat = material_any(rp, cfg, w);
    if (mode == FPT_DIAGNOSTIC_MATERIAL)
        return mk(clampf(mat.rgb.x, 0.f, 1.f), clampf(mat.rgb.y, 0.f, 1.f), clampf(mat.rgb.z, 0.f, 1.f));
    if (mode == FPT_DIAGNOSTIC_PATH_DIRECT) {
        f3 c = mk(0.f, 0.f, 0.f);
        if (cfg.sun[0] == 1.0f) {  // sunContribution (SM:5863-5867)
            int unused = 0;
            c = sun_contribution<SCENE_ANY>(rp, xy, 0.0f, material_any(rp, cfg, w), normal_any(rp, cfg, w), cfg, w,
                                            unused);
        }
        return diag_hdr(c, cfg);
    }
    if (mode == FPT_DIAGNOSTIC_PATH_ENVIRONMENT) return diag_hdr(environment(rd, cfg), cfg);
    if (mode == FPT_DIAGNOSTIC_PATH_THROUGHPUT) {
        const f3 n = normal_any(rp, cfg, w);
        const f3 next = normalize(vmix(reflect(rd, n), random_vector(n, xy, 0.0f), mat.roughness));
        const float al = clampf(dot(next, n) * 0.5f + 0.5f, 0.f, 1.f);
        const f3 c = mat.rgb * al;
        return mk(clampf(c.x, 0.f, 1.f), clampf(c.y, 0.f, 1.f), clampf(c.z, 0.f, 1.f));
    }
    return diag_normal(normal_any(rp, cfg, w));
}

__device__ __forceinline__ uchar4 encode_rgba8(f3 c) {
    return make_uchar4((unsigned char)(clampf(c.x, 0.f, 1.f) * 255.0f), (unsigned char)(clampf(c.y, 0.f, 1.f) * 255.0f),
                       (unsigned char)(clampf(c.z, 0.f, 1.f) * 255.0f), 255);
}

// sdf_diagnostic_kernel (SM:7375-7388): 16x16 pixel tiles, 8x8 per wave.
__global__ __launch_bounds__(256) void sdf_diagnostic_kernel(const FptRenderConfig *__restrict__ cfgp,
                                                             FptDiagnosticConfig dg, uchar4 *__restrict__ out) {
    const FptRenderConfig &cfg = *cfgp;
    uint32_t x, y;
    tile_pixel(x, y);
    if (x >= cfg.width || y >= cfg.height) return;
    Work w = {0, 0};
    const f3 c = sdf_diagnostic(screen_uv_any(x, y, cfg), cfg, dg, w);
    out[size_t(cfg.height - 1u - y) * cfg.width + x] = encode_rgba8(c);
}

hipError_t launch_sdf_diagnostic(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg,
                                 const FptDiagnosticConfig &dg, uchar4 *out, hipStream_t stream) {
    dim3 grid((hcfg.width + 15u) / 16u, (hcfg.height + 15u) / 16u, 1u);
    hipLaunchKernelGGL(sdf_diagnostic_kernel, grid, dim3(256), 0, stream, dcfg, dg, out);
    return hipGetLastError();
}

hipError_t launch_sdf_accumulate(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg, float4 *accum,
                                 uint32_t start, uint32_t count, DeviceProfile *prof, hipStream_t stream) {
    const dim3 grid = tile_grid(hcfg.width, hcfg.height), block(256);
    const int kind = hcfg.sdf_id == FPT_SDF_MANDELBULBER ? SCENE_MB
                     : has_lowered_tables(hcfg)           ? SCENE_LOWERED
                                                          : SCENE_SDF;
#define FPT_ACCUMULATE(K)                                                                                           \
    if (prof)                                                                                                       \
        hipLaunchKernelGGL((sdf_accumulate_kernel<true, K>), grid, block, 0, stream, dcfg, accum, start, count, prof); \
    else                                                                                                            \
        hipLaunchKernelGGL((sdf_accumulate_kernel<false, K>), grid, block, 0, stream, dcfg, accum, start, count,     \
                           nullptr)
    if (kind == SCENE_MB) {
        FPT_ACCUMULATE(SCENE_MB);
    } else if (kind == SCENE_LOWERED) {
        FPT_ACCUMULATE(SCENE_LOWERED);
    } else {
        FPT_ACCUMULATE(SCENE_SDF);
    }
#undef FPT_ACCUMULATE
    return hipGetLastError();
}

hipError_t launch_trace_primary(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg, FptRayHit *hits,
                                hipStream_t stream) {
    hipLaunchKernelGGL(trace_primary_kernel, tile_grid(hcfg.width, hcfg.height), dim3(256), 0, stream, dcfg, hits);
    return hipGetLastError();
}

hipError_t launch_pixel_cost(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg, float *cost,
                             hipStream_t stream) {
    const dim3 grid = tile_grid(hcfg.width, hcfg.height), block(256);
    if (hcfg.sdf_id == FPT_SDF_MANDELBULBER)
        hipLaunchKernelGGL(pixel_cost_kernel<SCENE_MB>, grid, block, 0, stream, dcfg, cost);
    else if (has_lowered_tables(hcfg))
        hipLaunchKernelGGL(pixel_cost_kernel<SCENE_LOWERED>, grid, block, 0, stream, dcfg, cost);
    else
        hipLaunchKernelGGL(pixel_cost_kernel<SCENE_SDF>, grid, block, 0, stream, dcfg, cost);
    return hipGetLastError();
}

// estimate_focus_distance_kernel (SM:5880-5884), as the host's focus
// prepass uses it (metal_bridge.mm:4827-4838, 5455-5481): the estimate is
// written into the device config, so every later kernel of the render reads
// it as the configured focus distance.  One thread per config; configs
// without depth of field or with a focus distance set are left as they are.
__global__ void focus_prepass_kernel(FptRenderConfig *cfgs, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    FptRenderConfig &cfg = cfgs[i];
    if (!(cfg.camera_dof > 0.0f) || cfg.focus_distance > 0.0f) return;
    Work w = {0, 0};
    const float f = focus_estimate_any(cfg, w);
    cfg.focus_distance = f;
}

hipError_t launch_focus_prepass(FptRenderConfig *dcfgs, uint32_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(focus_prepass_kernel, dim3((n + 63u) / 64u), dim3(64), 0, stream, dcfgs, n);
    return hipGetLastError();
}

hipError_t launch_field_sample(const FptRenderConfig *dcfg, const float4 *pts, size_t n,
                               FptMandelbulberFieldSample *out, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(field_sample_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream, dcfg, pts, n, out);
    return hipGetLastError();
}

hipError_t launch_typed_soa_validation(const FptRenderConfig *dcfg, uint32_t n, uint32_t *counts,
                                       hipStream_t stream) {
    hipLaunchKernelGGL(typed_soa_validation_kernel, dim3((n + 255u) / 256u), dim3(256), 0, stream, dcfg, n, counts);
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
