// Persistent, path-regenerating SDF path tracer for gfx950, for one scene or
// a batch of scenes in one launch.
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
// samples are done it writes the running mean and takes the next work slot
// from a global counter (one atomic per wave).  Slots are pixels in 8x8 tile
// order, scene after scene for a batch, so the lanes of a wave work on
// neighbouring pixels of one scene except at scene boundaries.  Samples of a
// pixel are processed in order by one lane, so the running mean
// mix(c, s, 1/(i+1)) and every RNG key are exactly those of the
// straightforward kernel and the images are bit-identical.
//
// MULTI = false: one scene, its config read through a kernel-argument pointer
// (wave-uniform scalar loads).  MULTI = true: every lane carries the pointer
// to its scene's config (vector loads, cached), scenes may differ in size,
// spp and program.
//
// BULB = true (single scene whose program is one Mandelbulb leaf): the
// march advances one orbit iteration per loop trip instead of one whole
// distance evaluation.  A lane whose orbit ends (escape, iteration limit or
// non-finite radius) finalises the distance, takes its march step and
// starts the next orbit in the same trip, while the other lanes keep
// iterating, so a wave no longer waits at every step for its longest orbit
// (orbits run 1-12 iterations; measured lane utilisation of the VALU was
// 25%).  Each lane executes exactly the operations of leaf_mandelbulb in the
// same order, so the image is bit-identical.
//
// LEARNED = true (single scene): march steps inside the proposer's training
// box may use the learned distance proposal q (learned.h) instead of the exact
// distance.  A march is in "far" mode from the camera, and again whenever an
// exact step saw |d| > resume_distance; in far mode, while q >
// switch_distance, the step is step_scale * q (no exact evaluation).  The
// first proposal at or below switch_distance hands the march to the exact
// distance; if that exact distance is negative the far step overshot into the
// surface and the march backs up to the point before it.  Far steps have their
// own budget (the march's step limit); when it runs out the march stays
// exact.  Hits, normals, shading and termination tests are the exact ones.
//
// LOWERED = true: configs whose program carries flat-union, canonical or
// typed structure-of-arrays tables march and shade through them
// (map_scene / normal_scene); the other instances leave that code out.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

#include "fpt_device.h"
#include "fpt_internal.h"
#include "learned.h"

namespace fpt {

enum : int { PH_IDLE = 0, PH_MARCH = 1, PH_EVENT = 2 };

struct PersistentCounters {
    int primary_steps, secondary_steps, shadow_steps, normal_evals, bounces;
};

struct LearnedCounters {
    int proposer_calls, proposals, far_steps, backtracks;
};

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

// Lane-occupancy counters of the counting pass (FptLaneStats): the first
// active lane of the region adds the region's count, so summing the lanes'
// counters at the end gives one count per wave pass.
enum : int {
    LS_TRIPS = 0, LS_MARCH_PASSES, LS_MARCH_LANES, LS_FINISH_PASSES, LS_FINISH_LANES, LS_EVENT_PASSES,
    LS_EVENT_LANES, LS_NORMAL_PASSES, LS_NORMAL_LANES, LS_IDLE_LANES, LS_PROPOSER_LANES, LS_COUNT = 12
};
__device__ __forceinline__ void lane_stat(unsigned long long *ls, int passes, int lanes, bool member) {
    const unsigned long long m = __ballot(member);
    const unsigned long long active = __ballot(true);
    if (m && int(lane_id()) == __ffsll((long long)active) - 1) {
        if (passes >= 0) ls[passes] += 1u;
        ls[lanes] += __popcll(m);
    }
}

// Wave-level fetch of work slots: one atomic per wave, consecutive indices to
// the requesting lanes in lane order.  Must be reached by every lane.
// The lowered tables' normal as a call of its own, outside the persistent
// kernel's flattened body: the 28-primitive union renders 2.7% faster
// (same box: 81.8 against 84.0 ms) with the same operations.
__device__ __noinline__ f3 normal_scene_call(f3 p, const FptRenderConfig &cfg, Work &w) {
    return normal_scene(p, cfg, w);
}

__device__ __forceinline__ uint32_t fetch_slot(bool want, uint32_t *counter) {
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

// Slot index within a scene -> (x, y): 8x8 tiles, handed out in `order`
// (a permutation of the row-major tiles) when given, else row-major.
__device__ __forceinline__ void pixel_of(uint32_t k, uint32_t w, const uint32_t *order, uint32_t &x, uint32_t &y) {
    const uint32_t tiles_x = (w + 7u) >> 3;
    const uint32_t tile = order ? order[k >> 6] : k >> 6, in = k & 63u;
    x = (tile % tiles_x) * 8u + (in & 7u);
    y = (tile / tiles_x) * 8u + (in >> 3);
}

__host__ __device__ __forceinline__ uint32_t slot_count(uint32_t w, uint32_t h) {
    return ((w + 7u) / 8u) * ((h + 7u) / 8u) * 64u;
}

// Position of the r-th (0-based) set bit of m; requires r < popcount(m).
__device__ __forceinline__ uint32_t select_bit(unsigned long long m, uint32_t r) {
    uint32_t pos = 0u, word = uint32_t(m);
    uint32_t c = uint32_t(__popc(word));
    if (r >= c) {
        r -= c;
        word = uint32_t(m >> 32);
        pos = 32u;
    }
#pragma unroll
    for (uint32_t half = 16u; half >= 1u; half >>= 1) {
        c = uint32_t(__popc(word & ((1u << half) - 1u)));
        if (r >= c) {
            r -= c;
            word >>= half;
            pos += half;
        }
    }
    return pos;
}

__device__ __forceinline__ uint32_t as_word(float v) { return __float_as_uint(v); }
__device__ __forceinline__ uint32_t as_word(int v) { return uint32_t(v); }
__device__ __forceinline__ uint32_t as_word(uint32_t v) { return v; }
__device__ __forceinline__ void from_word(float &v, uint32_t w) { v = __uint_as_float(w); }
__device__ __forceinline__ void from_word(int &v, uint32_t w) { v = int(w); }
__device__ __forceinline__ void from_word(uint32_t &v, uint32_t w) { v = w; }

// Stores four 32-bit lane values into one 16-byte LDS word and, when `read`,
// takes the word's previous contents in exchange.
template <class A, class B, class C, class D>
__device__ __forceinline__ void xchg4(uint4 *slot, bool read, A &a, B &b, C &c, D &d) {
    uint4 t;
    if (read) t = *slot;
    *slot = make_uint4(as_word(a), as_word(b), as_word(c), as_word(d));
    if (read) {
        from_word(a, t.x);
        from_word(b, t.y);
        from_word(c, t.z);
        from_word(d, t.w);
    }
}

// Orders LDS accesses of different lanes of one wave: the compiler keeps
// every access on its side of this point.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Minimum resident waves per SIMD the register allocation must allow (exact
// and learned variants); the compiler caps VGPRs at 512 / waves.
#ifndef FPT_PERSISTENT_WAVES
#define FPT_PERSISTENT_WAVES 2
#endif
#ifndef FPT_LEARNED_WAVES
#define FPT_LEARNED_WAVES 2
#endif
// Lanes that must want a learned proposal before the network runs for them.
#ifndef FPT_PROPOSER_BATCH
#define FPT_PROPOSER_BATCH 32
#endif
// Interleaved-orbit variant for a single Mandelbulb leaf (on/off), and its
// finished lanes per finalisation batch.
#ifndef FPT_INTERLEAVED_BULB
#define FPT_INTERLEAVED_BULB 1
#endif
#ifndef FPT_ORBIT_BATCH
#define FPT_ORBIT_BATCH 40
#endif
// Minimum resident waves per SIMD for imported Mandelbulber scenes (MB).
#ifndef FPT_MANDELBULBER_PERSISTENT_WAVES
#define FPT_MANDELBULBER_PERSISTENT_WAVES kMandelbulberPersistentWaves
#endif
// Path stash of the single-scene exact instances (on/off).
#ifndef FPT_PATH_STASH
#define FPT_PATH_STASH 1
#endif

// Uses the LDS path stash: exact-march instances (single scene or batch).
__host__ __device__ constexpr bool uses_path_stash(bool multi, bool learned, bool mb) {
    return FPT_PATH_STASH != 0 && !learned && !mb;
}

// MB = true (single scene): an imported Mandelbulber scene, with the
// Mandelbulber camera, march (threshold, clamped step, FP32 fixed-point and
// distance stops; the bracketing refinement runs at the march's event),
// material and normal (Shaders.metal:473-511, 975-995, 3627-3751, 3871-3960).
template <bool PROFILE, bool MULTI, bool LEARNED, bool BULB = false, bool MB = false, bool LOWERED = false>
// flatten: every device function the kernel reaches is inlined, so no call
// (with its stack traffic for live state) ever sits on the march or the
// per-hit path, however the inliner's size budget would decide.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(
    LEARNED ? FPT_LEARNED_WAVES : (MB ? FPT_MANDELBULBER_PERSISTENT_WAVES : FPT_PERSISTENT_WAVES)), flatten))
void sdf_persistent_kernel(const FptRenderConfig *__restrict__ cfgs,
                                                             const BatchScene *__restrict__ scenes, uint32_t nscenes,
                                                             float4 *__restrict__ accum, uint32_t start,
                                                             uint32_t count, uint32_t *__restrict__ counter,
                                                             uint32_t total_slots, int event_batch, int refill,
                                                             DeviceProfile *prof, ProposerDevice prop,
                                                             const uint32_t *__restrict__ tile_order,
                                                             SampleSplit split) {
    // Path stash: each wave keeps up to 64 further paths parked in LDS, one
    // slot per lane index (slot j's phase lives in lane j's `slot_phase`).
    // Lanes whose path waits for the per-hit pass (or is idle) exchange it
    // for a parked marching path once `refill` lanes are not marching, and
    // the per-hit pass first pulls parked waiting paths into the lanes that
    // are not waiting; it runs once `event_batch` of the wave's up to 128
    // paths wait, or fewer than `event_batch` march.  A path's whole state
    // moves with it, so every path executes the same operations as without
    // the stash and images are bit-identical.
    constexpr bool STASH = uses_path_stash(MULTI, LEARNED, MB);
    constexpr int SG = BULB ? 14 : 13;  // 16-byte words per parked path
    __shared__ uint4 stash_lds[STASH ? 4 * SG * 64 : 1];
    uint4 *const st = stash_lds + (STASH ? (threadIdx.x >> 6) * SG * 64 : 0);
    int slot_phase = PH_IDLE;

    const FptRenderConfig *cp = cfgs;  // the lane's scene (fixed when !MULTI)
    float4 *out = accum;               // the lane's scene accumulation buffer

    Work w = {0, 0};
    PersistentCounters pc = {0, 0, 0, 0, 0};
    LearnedCounters lc = {0, 0, 0, 0};
    unsigned long long ls[LS_COUNT] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    int pixels_done = 0;

    // Per-scene constants, refreshed when a lane takes a pixel.
    uint32_t W = 0, H = 0, s_end = 0, s_begin = 0;
    uint32_t scene = 0;  // MULTI: index of the lane's scene
    f3 cam = mk(0.f, 0.f, 0.f);
    float focal = 1.f, aa = 0.f, maxd = 0.f, maxd2 = 0.f, far2 = 0.f, min_dist = 0.f, adapt = 0.f;
    bool check_t = false, sun_on = false;
    int bounces = 0, ni_first = 0, ni_next = 0;
    f3 sun_dir0 = mk(0.f, 0.f, 1.f);
    auto load_scene = [&]() __attribute__((always_inline)) {
        const FptRenderConfig &cfg = *cp;
        const uint32_t total = cfg.samples < 1u ? 1u : (cfg.samples > 512u ? 512u : cfg.samples);
        s_begin = MULTI ? 0u : start;
        // MULTI: `count` caps the samples per pixel (0 = all; autotuning proxies).
        s_end = MULTI ? (count && count < total ? count : total) : (start + count < total ? start + count : total);
        cam = ld3(cfg.camera_position);
        focal = 1.0f / tanf(cfg.camera_fov / 2.0f * kPi / 180.0f);
        aa = 0.3f / fmaxf(float(W), float(H));
        maxd = cfg.render[4];
        maxd2 = maxd * maxd;
        const float far = maxd * 0.99f;
        far2 = far * far;
        min_dist = cfg.render[3];
        adapt = cfg.render[5];
        check_t = cfg.sdf_id == FPT_SDF_PROGRAM && cfg.program_material[5] > 0.0f;
        bounces = int(cfg.render[0]) < 8 ? int(cfg.render[0]) : 8;
        if (cfg.sdf_bounce_cap > 0u && int(cfg.sdf_bounce_cap) < bounces) bounces = int(cfg.sdf_bounce_cap);
        ni_first = int(cfg.render[1]);
        ni_next = int(cfg.render[1] / (cfg.render[5] * 2.0f + 1.0f));
        sun_on = cfg.sun[0] == 1.0f;
        sun_dir0 = rotate_camera(mk(0.f, 0.f, 1.f), cfg.sun[1] * kPi / 180.0f, cfg.sun[2] * kPi / 180.0f);
    };
    if (!MULTI) {
        W = cfgs->width;
        H = cfgs->height;
        load_scene();
    }

    // Sample groups (single scene only): a slot is a run of a pixel's samples.
    const bool SPLIT = !MULTI && !LEARNED && split.groups > 1u;

    // Lane state.
    int phase = PH_IDLE;
    uint32_t pix = 0, s = 0, s_stop = 0;  // s_stop: end of the slot's samples
    UV uv = {0.f, 0.f};
    f3 acc = mk(0.f, 0.f, 0.f);
    int bounce = 0, local_ni = 0;
    float side = 1.0f, sky_mask = 0.0f;
    f3 light = mk(0.f, 0.f, 0.f), color = mk(1.f, 1.f, 1.f), gcol = mk(0.f, 0.f, 0.f), dr = mk(0.f, 0.f, 0.f);
    f3 hit_p = mk(0.f, 0.f, 0.f), hit_n = mk(0.f, 0.f, 0.f);
    Material mat;
    mat.rgb = mk(1.f, 1.f, 1.f);
    mat.roughness = 1.f; mat.specular = 0.f; mat.translucency = 0.f; mat.ior = 1.5f; mat.emission = 0.f;
    f3 p = mk(0.f, 0.f, 0.f), mdir = mk(0.f, 0.f, 1.f), mstart = mk(0.f, 0.f, 0.f);
    int it = 0, max_iter = 0;
    bool shadow = false;
    // Learned stepping state (LEARNED only).
    bool lfar = false, lastfar = false;
    int fit = 0;
    f3 psafe = mk(0.f, 0.f, 0.f);
    // Interleaved orbit state (BULB only): z, |z|, dz and the iteration count
    // of the distance evaluation in progress at p.
    v3<float> bz = {0.f, 0.f, 0.f};
    float br = 0.f, bdr = 1.f;
    int boi = 0;
    bool borb = false;
    float bpower = 8.f, bbail = 2.f;
    int biters = 0, bip = 8;
    if (BULB) {
        const float *bd = cfgs->sdf_program[0].data;
        bpower = bd[0];
        biters = int(bd[1]);
        bbail = bd[2];
        const int ipw = int(bpower);
        bip = (float(ipw) == bpower && ipw >= 1 && ipw <= 32) ? ipw : 0;
    }
    // Mandelbulber march state (MB only): the last distance, threshold and
    // step, whether the threshold band was reached, and a camera sample
    // outside a cut fish-eye (the path is black).
    bool mfound = false, invisible = false;
    float mdist = 0.f, mthr = 0.f, mstep = 0.f;
    bool retry = false;  // slot was tile padding: take another

    // sdfMarchIterationLimit (SM:3626-3633).
    auto march_limit = [&](int ni) __attribute__((always_inline)) {
        const int cap = MB ? 10000 : 360;
        return ni < cap ? ni : cap;
    };

    // Start the camera path of sample `s` of the current pixel.
    auto start_path = [&]() __attribute__((always_inline)) {
        const FptRenderConfig &cfg = *cp;
        const float frame = float(s);
        float jx, jy;
        random_point(aa, uv, frame, jx, jy);
        if (MB) {
            invisible = !mb_projection_visible(uv.x + jx, uv.y + jy, cfg);
            dr = invisible ? mk(0.f, 0.f, 1.f) : mb_camera_ray(uv.x + jx, uv.y + jy, cfg);
        } else {
            dr = rotate_camera_roll(normalize(mk(uv.x + jx, uv.y + jy, focal)), cfg.camera_yaw_pitch[0],
                                    cfg.camera_yaw_pitch[1], cfg.camera_roll);
        }
        p = cam;
        if (!MB || !invisible) apply_dof(uv, frame, cfg, p, dr);
        const float gt = clampf((dr.y + 1.0f) / 2.0f, 0.f, 1.f);
        gcol = vmix(ld3(&cfg.background_gradient[3]), ld3(&cfg.background_gradient[0]), gt);
        bounce = 0;
        local_ni = ni_first;
        side = 1.0f;
        sky_mask = 0.0f;
        light = mk(0.f, 0.f, 0.f);
        color = mk(1.f, 1.f, 1.f);
        mstart = p;
        mdir = dr;
        it = 0;
        max_iter = march_limit(local_ni);
        shadow = false;
        lfar = LEARNED;
        lastfar = false;
        fit = 0;
        mfound = false;
        mstep = 0.f;  // a march that starts inside the band refines in place
        phase = MB && invisible ? PH_EVENT : PH_MARCH;
    };

    // Take a work slot (or go idle).  Wave-uniform call.
    auto take_slot = [&](bool want) __attribute__((always_inline)) {
        const uint32_t k = fetch_slot(want, counter);
        retry = false;
        if (!want) return;
        if (k >= total_slots) {
            phase = PH_IDLE;
            return;
        }
        uint32_t local = k, run = 0u;
        if (SPLIT) {
            // Slots run tile by tile, each tile's sample runs one after the
            // other: k = ((tile rank * groups + run) * 64 + pixel in tile).
            const uint32_t kg = k >> 6, rank = kg / split.groups;
            run = kg - rank * split.groups;
            local = (rank << 6) + (k & 63u);
        }
        if (MULTI) {
            // Slot in the hand-out order: tile_order permutes the batch's global
            // 8x8 tiles (scene after scene, row-major within a scene).
            const uint32_t g = tile_order ? tile_order[k >> 6] * 64u + (k & 63u) : k;
            // Binary search of the scene owning slot k.
            uint32_t lo = 0, hi = nscenes - 1u;
            while (lo < hi) {
                uint32_t mid = (lo + hi + 1u) >> 1;
                if (scenes[mid].slot_begin <= g) lo = mid;
                else hi = mid - 1u;
            }
            const BatchScene sc = scenes[lo];
            scene = lo;
            cp = cfgs + lo;
            out = accum + sc.accum_offset;
            W = sc.width;
            H = sc.height;
            local = g - sc.slot_begin;
            load_scene();
        }
        uint32_t x, y;
        pixel_of(local, W, MULTI ? nullptr : tile_order, x, y);
        if (x >= W || y >= H) {
            retry = true;
            return;
        }
        pix = y * W + x;
        uv = MB ? screen_uv_any<SCENE_MB>(x, y, *cp) : screen_uv(x, y, W, H);
        s = SPLIT ? s_begin + run * split.chunk : s_begin;
        s_stop = SPLIT && s + split.chunk < s_end ? s + split.chunk : s_end;
        if (!SPLIT && s_begin != 0u) {
            float4 prev = out[pix];
            acc = mk(prev.x, prev.y, prev.z);
        } else {
            acc = mk(0.f, 0.f, 0.f);
        }
        start_path();
    };

    // Moves the lane's path into stash slot j; with `read`, the path parked
    // there comes back in exchange (32-bit fields in 16-byte words, slot-
    // major so the 64 lanes of a wave touch 64 consecutive words).
    auto move_path = [&](uint32_t j, bool read) __attribute__((always_inline)) {
        uint4 *q = st + j;
        uint32_t flags = uint32_t(shadow) | (uint32_t(borb) << 1) | (s_stop << 16);
        xchg4(q + 0 * 64, read, phase, pix, s, bounce);
        xchg4(q + 1 * 64, read, uv.x, uv.y, local_ni, it);
        xchg4(q + 2 * 64, read, acc.x, acc.y, acc.z, side);
        xchg4(q + 3 * 64, read, light.x, light.y, light.z, sky_mask);
        xchg4(q + 4 * 64, read, color.x, color.y, color.z, max_iter);
        xchg4(q + 5 * 64, read, gcol.x, gcol.y, gcol.z, flags);
        xchg4(q + 6 * 64, read, dr.x, dr.y, dr.z, mat.roughness);
        xchg4(q + 7 * 64, read, hit_p.x, hit_p.y, hit_p.z, mat.specular);
        xchg4(q + 8 * 64, read, hit_n.x, hit_n.y, hit_n.z, mat.translucency);
        xchg4(q + 9 * 64, read, mat.rgb.x, mat.rgb.y, mat.rgb.z, mat.ior);
        xchg4(q + 10 * 64, read, p.x, p.y, p.z, mat.emission);
        if (BULB) {
            xchg4(q + 11 * 64, read, mdir.x, mdir.y, mdir.z, bdr);
            xchg4(q + 12 * 64, read, mstart.x, mstart.y, mstart.z, boi);
            xchg4(q + 13 * 64, read, bz.x, bz.y, bz.z, br);
        } else {
            uint32_t pad1 = 0u;
            xchg4(q + 11 * 64, read, mdir.x, mdir.y, mdir.z, scene);
            xchg4(q + 12 * 64, read, mstart.x, mstart.y, mstart.z, pad1);
        }
        shadow = (flags & 1u) != 0u;
        borb = (flags & 2u) != 0u;
        s_stop = flags >> 16;
    };
    // The r-th lane (in lane order) with `want` exchanges its path with the
    // path in the r-th slot of `offer`, for as many pairs as both sides have.
    // Slots are counted from `rot`, one past the last slot taken, so parked
    // paths are taken round-robin and none waits behind lower-numbered slots.
    // Wave-uniform call.
    uint32_t rot = 0u;
    auto exchange_in = [&](bool want, unsigned long long offer) __attribute__((always_inline)) {
        const unsigned long long wm = __ballot(want);
        const uint32_t nw = uint32_t(__popcll(wm)), no = uint32_t(__popcll(offer));
        const uint32_t n = nw < no ? nw : no;
        if (n == 0u) return;
        const uint32_t r = uint32_t(__popcll(wm & ((1ull << lane_id()) - 1ull)));
        const unsigned long long off = rot ? (offer >> rot) | (offer << (64u - rot)) : offer;
        wave_lds_sync();
        if (want && r < n) {
            move_path((select_bit(off, r) + rot) & 63u, true);
            if (MULTI) {  // the path's scene: its config, image and constants
                const BatchScene sc = scenes[scene];
                cp = cfgs + scene;
                out = accum + sc.accum_offset;
                W = sc.width;
                H = sc.height;
                load_scene();
            }
        }
        wave_lds_sync();
        slot_phase = int(st[lane_id()].x);
        rot = (rot + select_bit(off, n - 1u) + 1u) & 63u;
    };

    take_slot(true);
    while (__any(retry)) take_slot(retry);
    if (STASH) {
        // Park the first paths and take a second pixel per lane.
        move_path(lane_id(), false);
        slot_phase = phase;
        phase = PH_IDLE;
        take_slot(true);
        while (__any(retry)) take_slot(retry);
    }

    while (true) {
        if (PROFILE) {
            if (lane_id() == 0u) ls[LS_TRIPS] += 1u;
            lane_stat(ls, -1, LS_IDLE_LANES, phase == PH_IDLE);
        }
        float q = 0.f;
        bool use_q = false, wait_q = false;
        if (LEARNED) {
            // The network is wave-collective (every lane runs it), so it is
            // evaluated for a batch: once FPT_PROPOSER_BATCH lanes want a
            // proposal, or when no marching lane has an exact step to take
            // meanwhile.  Until then the lanes that want one wait.
            const bool want = phase == PH_MARCH && lfar && proposer_in_box(prop, p);
            const unsigned long long wm = __ballot(want);
            const unsigned long long exact_m = __ballot(phase == PH_MARCH && !want);
            if (wm && (__popcll(wm) >= FPT_PROPOSER_BATCH || !exact_m)) {
                use_q = want;
                q = proposer_eval(prop, p);
                if (PROFILE) {
                    lc.proposer_calls += lane_id() == 0u;
                    lc.proposals += use_q;
                    lane_stat(ls, -1, LS_PROPOSER_LANES, use_q);
                }
            } else {
                wait_q = want;
            }
        }
        if (BULB) {
            // Orbit iterations until at least FPT_ORBIT_BATCH lanes of the
            // wave have finished their distance evaluation (or no orbit is
            // left running); lanes that finish early wait, so the
            // finalisation and march step below run for a batch of lanes.
            const bool act = phase == PH_MARCH;
            if (act && !borb) {  // start the distance evaluation at p (leaf_mandelbulb)
                bz = v3<float>{p.x, p.y, p.z};
                br = orbit_radius(bz);
                bdr = 1.0f;
                boi = 0;
                borb = true;
                w.distance_evals++;
            }
            bool fin = act && boi >= biters;
            while (true) {
                const bool run = act && !fin;
                if (!__any(run)) break;
                if (PROFILE) lane_stat(ls, LS_MARCH_PASSES, LS_MARCH_LANES, run);
                if (run) {
                    const v3<float> old = bz;
                    if (bip == 8) bulb_step<float, 8>(bz, br, bdr, bpower, 8);
                    else bulb_step<float, 0>(bz, br, bdr, bpower, bip);
                    bz = add(bz, v3<float>{p.x, p.y, p.z});
                    br = orbit_radius(bz);
                    boi++;
                    if (!isfinite(br)) {
                        bz = old;
                        br = orbit_radius(bz);
                        fin = true;
                    } else {
                        fin = br > bbail || boi >= biters;
                    }
                }
                if (__popcll(__ballot(fin)) >= FPT_ORBIT_BATCH) break;
            }
            const FptRenderConfig &cfg = *cp;
            if (PROFILE) lane_stat(ls, LS_FINISH_PASSES, LS_FINISH_LANES, fin);
            if (fin) {
                borb = false;
                w.orbit_iterations += boi;
                float dd = bulb_finish(br, bdr);
                dd = dd / fmaxf(1.0f, 1.0e-6f);  // the program's distance scale (one leaf, no transform)
                const float d = mixf(dd, dd / 2.0f, cfg.render[6]);  // mapSdf (SM:3612-3615)
                ++it;
                bool stop = !isfinite(d);
                if (!stop) {
                    const float o = fabsf(d) * 0.99f;
                    p = p + mdir * o;
                    const f3 df = mstart - p;  // lod_falloff 0.0002 (SM:3771-3776)
                    float lod = mixf(min_dist, mixf(min_dist, 0.1f, dot(df, df) * 0.0002f), adapt);
                    if (check_t && user_material(p, cfg, w).translucency > 0.0f) lod = 0.0001f;
                    stop = o < lod || maxd < o || it >= max_iter;
                }
                if (stop) phase = PH_EVENT;
            }
        } else if (MB && phase == PH_MARCH) {
            // marchMandelbulber's loop (SM:3711-3733), one step per trip.
            if (PROFILE) lane_stat(ls, LS_MARCH_PASSES, LS_MARCH_LANES, true);
            const FptRenderConfig &cfg = *cp;
            bool stop = it >= max_iter;
            if (!stop) {
                mthr = mb_threshold(p, cfg);
                mdist = map_kifs(p, cfg, w);
                ++it;
                if (!isfinite(mdist)) {
                    stop = true;
                } else if (mdist < mthr) {
                    mfound = true;
                    stop = true;
                } else {
                    mstep = mb_step(mdist, mthr, cfg);
                    const f3 nx = p + mdir * mstep;
                    if (same3(nx, p)) {
                        stop = true;
                    } else {
                        p = nx;
                        stop = length(p - mstart) >= maxd;
                    }
                }
            }
            if (stop) phase = PH_EVENT;
        } else if (!BULB && phase == PH_MARCH && !(LEARNED && wait_q)) {
            if (PROFILE) lane_stat(ls, LS_MARCH_PASSES, LS_MARCH_LANES, true);
            const FptRenderConfig &cfg = *cp;
            if (LEARNED && use_q && q > prop.switch_distance) {
                // Far step on the proposal.
                psafe = p;
                p = p + mdir * (q * prop.step_scale);
                lastfar = true;
                if (++fit >= max_iter) lfar = false;
                if (PROFILE) lc.far_steps++;
            } else {
                if (LEARNED && use_q) lfar = false;  // within switch_distance: exact from here
                const float d = LOWERED ? map_scene(p, cfg, w) : map_sdf(p, cfg, w);
                ++it;
                if (LEARNED && lastfar && use_q && d < 0.0f) {
                    // The last far step crossed the surface: back up.
                    p = psafe;
                    lastfar = false;
                    if (PROFILE) lc.backtracks++;
                    if (it >= max_iter) phase = PH_EVENT;
                } else {
                    lastfar = false;
                    bool stop = !isfinite(d);
                    if (!stop) {
                        const float o = fabsf(d) * 0.99f;
                        p = p + mdir * o;
                        const f3 df = mstart - p;  // lod_falloff 0.0002 (SM:3771-3776)
                        float lod = mixf(min_dist, mixf(min_dist, 0.1f, dot(df, df) * 0.0002f), adapt);
                        if (check_t && user_material(p, cfg, w).translucency > 0.0f) lod = 0.0001f;
                        stop = o < lod || maxd < o || it >= max_iter;
                        if (LEARNED && fabsf(d) > prop.resume_distance && fit < max_iter) lfar = true;
                    }
                    if (stop) phase = PH_EVENT;
                }
            }
        }
        const unsigned long long ev = __ballot(phase == PH_EVENT);
        const unsigned long long mar = __ballot(phase == PH_MARCH);
        if (STASH) {
            const unsigned long long sev = __ballot(slot_phase == PH_EVENT);
            const unsigned long long smar = __ballot(slot_phase == PH_MARCH);
            if (!ev && !mar && !sev && !smar) break;
            // The per-hit pass runs once `event_batch` paths wait, or once at
            // least as many paths wait as march: when the wave's pixels run
            // out and fewer paths are left, a threshold on waiting paths
            // alone would hold every waiting path until the slowest march
            // ends.
            const int nev = __popcll(ev) + __popcll(sev);
            const int nmar = __popcll(mar) + __popcll(smar);
            if (nev >= event_batch || (nev > 0 && nev >= nmar)) {
                // Per-hit pass: waiting paths from the stash into the lanes
                // that are not waiting.
                if (sev && ~ev) exchange_in(phase != PH_EVENT, sev);
            } else {
                // Keep marching: parked marching paths into the lanes that
                // wait or idle, once enough of them do.
                if (smar && 64 - __popcll(mar) >= refill) exchange_in(phase != PH_MARCH, smar);
                continue;
            }
        } else {
            if (!ev && !mar) break;
            if (!(__popcll(ev) >= event_batch || (ev && !mar))) continue;
        }

        bool need_slot = false;
        if (PROFILE) lane_stat(ls, LS_EVENT_PASSES, LS_EVENT_LANES, phase == PH_EVENT);
        if (phase == PH_EVENT) {
            const FptRenderConfig &cfg = *cp;
            const float frame = float(s);
            bool path_done = false;
            bool do_scatter = false;
            if (MB && mfound) {
                // marchMandelbulber's bracketing refinement (SM:3737-3750):
                // up to 30 half-steps towards (threshold * (1 - 0.001
                // detail), threshold), with the last march threshold.
                const float limit = 1.0f - 0.001f * fmaxf(cfg.vset_values[129], 0.0f);
                float step = mstep * 0.5f, d = mdist;
                for (int r = 0; r < 30; ++r) {
                    if (d < mthr && d > mthr * limit) break;
                    if (d > mthr) {
                        const f3 nx = p + mdir * step;
                        if (same3(nx, p)) break;
                        p = nx;
                    } else if (d < mthr * limit) {
                        const f3 nx = p - mdir * step;
                        if (same3(nx, p)) break;
                        p = nx;
                    }
                    d = map_kifs(p, cfg, w);
                    step *= 0.5f;
                }
                mfound = false;
            }
            if (MB && invisible) {
                // Outside a cut fish-eye: renderPath returns black (SM:5938-5941).
                invisible = false;
                path_done = true;
            } else if (shadow) {
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
                    if (PROFILE) lane_stat(ls, LS_NORMAL_PASSES, LS_NORMAL_LANES, true);
                    mat = MB ? material_any<SCENE_MB>(p, cfg, w) : user_material(p, cfg, w);
                    if (mat.emission > 0.001f) light = light + mat.rgb * mat.emission;
                    hit_n = MB ? normal_mandelbulber(p, cfg, w) : (LOWERED ? normal_scene_call(p, cfg, w) : normal_at(p, cfg, w));
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
                        max_iter = march_limit(ni_first);
                        mstep = 0.f;
                        shadow = true;
                        lfar = false;
                        lastfar = false;
                        fit = 0;
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
                        color = color * (vmix(mk(1.f, 1.f, 1.f), mat.rgb, 0.35f) * (1.0f - fres * 0.5f));
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
                    max_iter = march_limit(local_ni);
                    mstep = 0.f;
                    lfar = false;
                    lastfar = false;
                    fit = 0;
                    phase = PH_MARCH;
                }
            }
            if (path_done) {
                f3 col = light * color;
                if (cfg.world[6] == 1.0f) col = vmix(col, gcol, sky_mask);
                col = vmin(col, 8.0f);
                if (SPLIT) {
                    // Agent-scope stores write through the L2, so the fence at
                    // a run's end finds no dirty colour lines to write back:
                    // one 8-byte store for the aligned pair, one 4-byte store.
                    const size_t e = 3u * (size_t(pix) * split.stride + s);
                    float *cs = split.colors + e;
                    const bool even = (e & 1u) == 0u;
                    const float lone = even ? col.z : col.x;
                    const unsigned long long pair =
                        even ? (uint64_t(__float_as_uint(col.y)) << 32) | __float_as_uint(col.x)
                             : (uint64_t(__float_as_uint(col.z)) << 32) | __float_as_uint(col.y);
                    __hip_atomic_store(reinterpret_cast<unsigned long long *>(even ? cs : cs + 1), pair,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(even ? cs + 2 : cs, lone, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                else acc = vmix(acc, col, 1.0f / float(s + 1u));
                ++s;
                if (s < s_stop) {
                    start_path();
                } else {
                    if (SPLIT) {
                        // The last run of the pixel to finish folds every
                        // sample's colour in order: a release fence before
                        // the (relaxed, device-scope) count publishes this
                        // run's colours, an acquire fence after it, in the
                        // last run only, makes the others' visible.
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                        if (atomicAdd(&split.done[pix], 1u) + 1u == split.groups) {
                            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                            f3 a = mk(0.f, 0.f, 0.f);
                            if (s_begin != 0u) {
                                const float4 prev = out[pix];
                                a = mk(prev.x, prev.y, prev.z);
                            }
                            const float *cs = split.colors + 3u * size_t(pix) * split.stride;
                            for (uint32_t i = s_begin; i < s_end; ++i)
                                a = vmix(a, mk(cs[3u * i], cs[3u * i + 1u], cs[3u * i + 2u]), 1.0f / float(i + 1u));
                            out[pix] = make_float4(a.x, a.y, a.z, 1.0f);
                            ++pixels_done;
                        }
                    } else {
                        out[pix] = make_float4(acc.x, acc.y, acc.z, 1.0f);
                        ++pixels_done;
                    }
                    need_slot = true;
                }
            }
        }
        // Wave-uniform hand-out for lanes whose pixel is finished.
        take_slot(need_slot);
        while (__any(retry)) take_slot(retry);
        if (STASH) {
            // Lanes left idle (no pixel to take) take parked marching paths:
            // once the wave's pixels run out, most trips end here, and the
            // refill above would not run.
            const unsigned long long smar = __ballot(slot_phase == PH_MARCH);
            if (smar && 64 - __popcll(__ballot(phase == PH_MARCH)) >= refill) exchange_in(phase != PH_MARCH, smar);
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
        for (int k = 0; k < LS_COUNT; ++k) wadd(&prof->lane[k], ls[k]);
        if (LEARNED) {
            wadd(&prof->proposer_calls, (unsigned long long)lc.proposer_calls);
            wadd(&prof->proposals, (unsigned long long)lc.proposals);
            wadd(&prof->far_steps, (unsigned long long)lc.far_steps);
            wadd(&prof->backtracks, (unsigned long long)lc.backtracks);
        }
    }
}

// Counter reset as a kernel (not a memset node), so a captured graph replays
// it.  Internal linkage: every translation unit that launches it has its own.
static __global__ void zero_counter_kernel(uint32_t *c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0u;
}

// Clears n words (the sample-run counters of a frame) with a kernel.
static __global__ void zero_words_kernel(uint32_t *p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

}  // namespace fpt
