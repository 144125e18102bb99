// Internal interfaces between the HIP kernels and the C-ABI layer.
#pragma once

#ifndef __HIPCC_RTC__  // the runtime compiler provides these itself
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#endif

#include "../../include/fpt_abi.h"

namespace fpt {

// Device-side work counters (one render, all pixels), filled by the
// profiling variant of the render kernel.  Categories follow
// Shaders.metal:7650-7704.
struct DeviceProfile {
    unsigned long long distance_evals;
    unsigned long long orbit_iterations;
    unsigned long long primary_steps;
    unsigned long long secondary_steps;
    unsigned long long shadow_steps;
    unsigned long long normal_evals;
    unsigned long long bounces;
    unsigned long long pixels;
    // Learned stepping (zero otherwise): wave-collective proposer
    // evaluations, lanes that used a proposal, far steps taken, backtracks.
    unsigned long long proposer_calls;
    unsigned long long proposals;
    unsigned long long far_steps;
    unsigned long long backtracks;
    // Lane occupancy of the persistent kernel (FptLaneStats).
    unsigned long long lane[12];
};

struct ProposerDevice;

// Sample groups of the single-scene persistent kernels: each pixel's samples
// [start, start + count) are split into `groups` consecutive runs of `chunk`
// samples, each a work slot of its own, so different lanes render them at
// once.  Every sample's clamped colour goes to colors[3 (pix * stride + s)]
// (three floats); the run that finishes last (done[pix] counts finished
// runs) folds the colours in sample order into the running mean, exactly as
// one lane would have.  groups <= 1: one slot per pixel, the running mean in
// a register.
struct SampleSplit {
    float *colors;
    uint32_t *done;
    uint32_t groups, chunk, stride;
};

// One scene of a batched launch: its work slots are
// [slot_begin, slot_begin + slots) of the global slot range (8x8 tiles of its
// pixels), its image is accum[accum_offset .. accum_offset + width*height).
struct BatchScene {
    uint32_t slot_begin;
    uint32_t slots;
    uint32_t width;
    uint32_t height;
    unsigned long long accum_offset;
};

// Waves per SIMD of the persistent instance for imported Mandelbulber
// scenes (no path stash, so no LDS limit): its launch defaults to this many
// workgroups per compute unit.  ifs-20.fract at 1080p, 8 spp: 910 ms at 2,
// 674 at 3, 580 at 4, 543 at 5, 530 at 6, 521 at 7, 508 at 8, against
// 617 ms for the one-path-per-lane kernel (images identical).
constexpr int kMandelbulberPersistentWaves = 8;

#ifndef __HIPCC_RTC__  // host launchers
hipError_t launch_zero_counter(uint32_t *counter, hipStream_t stream);
uint32_t persistent_slots(uint32_t width, uint32_t height);
// Clears n words on the stream with a kernel (capturable).
hipError_t launch_zero_words(uint32_t *p, uint32_t n, hipStream_t stream);
// The scene's program is one Mandelbulb leaf: the persistent kernel's
// interleaved-orbit variant applies.
bool single_bulb_program(const FptRenderConfig &c);
// The single-scene persistent instance that renders c (learned: with a
// step proposer) keeps parked paths in LDS (the path stash).
bool persistent_uses_stash(const FptRenderConfig &c, bool learned);
// The batched (multi-scene) persistent instance keeps parked paths in LDS.
bool batch_uses_stash();
// Event batch of a persistent launch given 0, counted over a wave's lanes
// and parked paths with the path stash: 64 for launches in sample runs and
// for batches of scenes (with the rule that also runs the per-hit work once
// as many paths wait as march, larger batches change nothing), 48 for one
// slot per pixel; 24 lanes without the stash.
int default_event_batch(bool stash, bool runs);
// Event batch of the imported-Mandelbulber instance (no stash, 8 waves per
// SIMD; ifs-20 at 1080p, 64 spp in sample runs: 3,634 ms at 24, 3,269 at
// 32, 3,144 at 40, 3,190 at 48, 3,484 at 56).
constexpr int kMandelbulberEventBatch = 40;
// Event batch of the learned-step instance (no stash; 1080p Mandelbulb,
// 64 spp: 2,386 ms at 8, 2,274 at 16, 2,303 at 20, 2,378 at 24, 2,855 at
// 32).
constexpr int kLearnedEventBatch = 16;
// Non-marching lanes of a wave before they take parked marching paths: 6
// for single-scene plans (the 1080p Mandelbulb in 8 sample runs: 551 ms at
// 4 runs and refill 4, 549 at 8 runs, 546 at 8 runs and refill 6), 4 for
// batches of scenes (64-scene batch: 747 ms at 4, 752 at 6, 760 at 8).
constexpr int kDefaultRefill = 6;
constexpr int kBatchRefill = 4;
// The scene's program carries flat-union, canonical or typed-SoA tables: the
// kernels' instances with the table evaluators apply.
bool has_lowered_tables(const FptRenderConfig &c);
// tile_order (optional, device): hand-out order of the batch's global 8x8
// tiles (scene after scene, row-major within a scene).  event_batch 0 takes
// the instance's default.
hipError_t launch_sdf_batch(const FptRenderConfig *dcfgs, const BatchScene *dscenes, uint32_t nscenes,
                            uint32_t total_slots, float4 *accum, uint32_t *counter, uint32_t workgroups,
                            int event_batch, uint32_t spp_cap, DeviceProfile *prof, bool lowered,
                            const uint32_t *tile_order, hipStream_t stream);

hipError_t launch_sdf_accumulate(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg, float4 *accum,
                                 uint32_t start, uint32_t count, DeviceProfile *prof, hipStream_t stream);
// Scene front end (scene_front.hip, fract_import.hip).
void scene_default_config(FptRenderConfig &c);
bool fract_load(const std::string &path, const FptSceneOptions &opt, FptRenderConfig &c, std::string &output_name,
                char *error, size_t error_len);
// HDRI hydration of a host config (hdri.hip): clears the LUT, then, for
// world mode 2 with HDRI enabled, fills it from the Radiance file at
// hdri_path.  Returns false with a message on failure.
bool hydrate_hdri(FptRenderConfig &cfg, char *error, size_t error_len);
uint16_t half_bits_round_up(float v);
// Writes the focus estimate into each of the n device configs that has
// depth of field and no focus distance.
hipError_t launch_focus_prepass(FptRenderConfig *dcfgs, uint32_t n, hipStream_t stream);
hipError_t launch_trace_primary(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg, FptRayHit *hits,
                                hipStream_t stream);
// Work of one whole path sample per pixel into cost[W*H] (row-major).
hipError_t launch_pixel_cost(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg, float *cost,
                             hipStream_t stream);
hipError_t launch_field_sample(const FptRenderConfig *dcfg, const float4 *pts, size_t n,
                               FptMandelbulberFieldSample *out, hipStream_t stream);
hipError_t launch_surface_sample(const FptRenderConfig *dcfg, const float4 *pts, size_t n, float4 *out,
                                 hipStream_t stream);
// counts: {distance failures, gradient failures, max distance error bits,
// max gradient error bits}, zeroed by the caller.
hipError_t launch_typed_soa_validation(const FptRenderConfig *dcfg, uint32_t n, uint32_t *counts,
                                       hipStream_t stream);
// tile_order (optional, device): the order in which the 8x8 tiles are handed
// out (a permutation of the row-major tile indices).  event_batch 0 takes
// default_event_batch of the instance launched.
// split: sample runs of the single-scene instances (SampleSplit; groups <= 1
// renders one slot per pixel); the launcher clears split.done.
hipError_t launch_sdf_persistent(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg, float4 *accum,
                                 uint32_t start, uint32_t count, uint32_t *counter, uint32_t workgroups,
                                 int event_batch, int refill, DeviceProfile *prof, const ProposerDevice *prop,
                                 const uint32_t *tile_order, const SampleSplit &split, hipStream_t stream);
hipError_t launch_sdf_diagnostic(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg,
                                 const FptDiagnosticConfig &dg, uchar4 *out, hipStream_t stream);
hipError_t launch_present(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg, const float4 *accum,
                          uchar4 *out, hipStream_t stream);

#endif

}  // namespace fpt
