// Host launchers of the persistent SDF path tracer (kernel in
// persistent_kernel.h, shared with the runtime-specialised build).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "persistent_kernel.h"

namespace fpt {

hipError_t launch_zero_counter(uint32_t *counter, hipStream_t stream) {
    hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(64), 0, stream, counter);
    return hipGetLastError();
}

uint32_t persistent_slots(uint32_t width, uint32_t height) { return slot_count(width, height); }

hipError_t launch_zero_words(uint32_t *p, uint32_t n, hipStream_t stream) {
    if (n) hipLaunchKernelGGL(zero_words_kernel, dim3((n + 255u) / 256u), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

bool single_bulb_program(const FptRenderConfig &c) {
    return FPT_INTERLEAVED_BULB && c.sdf_id == FPT_SDF_PROGRAM && c.sdf_program_count == 1u &&
           c.sdf_program[0].opcode == FPT_SDF_OP_MANDELBULB && !has_lowered_tables(c);
}

bool persistent_uses_stash(const FptRenderConfig &c, bool learned) {
    return uses_path_stash(false, learned, c.sdf_id == FPT_SDF_MANDELBULBER);
}

int default_event_batch(bool stash, bool runs) { return stash ? (runs ? 64 : 48) : 24; }

bool batch_uses_stash() { return uses_path_stash(true, false, false); }

// The dispatch condition of SM:3576-3584, computed exactly as the device's
// has_lowered(): each typed count capped at the table size before the sum,
// so a huge count cannot wrap the sum to zero on one side only.
bool has_lowered_tables(const FptRenderConfig &c) {
    if (c.sdf_id != FPT_SDF_PROGRAM || c.sdf_program_count == 0u) return false;
    const FptTypedSoAProgram &t = c.sdf_typed_soa;
    const uint32_t cap = uint32_t(FPT_SDF_FLAT_UNION_MAX_PRIMITIVES);
    const uint32_t typed = std::min(t.sphere_count, cap) + std::min(t.box_count, cap) + std::min(t.plane_count, cap);
    return typed > 0u || c.sdf_canonical_count > 0u || c.sdf_flat_union_count > 0u;
}

namespace {

template <bool PROFILE, bool MULTI, bool LEARNED, bool BULB, bool MB, bool LOWERED>
void launch_instance(dim3 grid, hipStream_t stream, const FptRenderConfig *cfgs, const BatchScene *scenes,
                     uint32_t nscenes, float4 *accum, uint32_t start, uint32_t count, uint32_t *counter,
                     uint32_t slots, int event_batch, int refill, DeviceProfile *prof, const ProposerDevice &prop,
                     const uint32_t *order, const SampleSplit &split) {
    hipLaunchKernelGGL((sdf_persistent_kernel<PROFILE, MULTI, LEARNED, BULB, MB, LOWERED>), grid, dim3(256), 0,
                       stream, cfgs, scenes, nscenes, accum, start, count, counter, slots, event_batch, refill,
                       PROFILE ? prof : nullptr, prop, order, split);
}

template <bool MULTI, bool LEARNED, bool BULB, bool MB, bool LOWERED>
void launch_profiled(bool profiled, dim3 grid, hipStream_t stream, const FptRenderConfig *cfgs,
                     const BatchScene *scenes, uint32_t nscenes, float4 *accum, uint32_t start, uint32_t count,
                     uint32_t *counter, uint32_t slots, int event_batch, int refill, DeviceProfile *prof,
                     const ProposerDevice &prop, const uint32_t *order, const SampleSplit &split) {
    if (profiled)
        launch_instance<true, MULTI, LEARNED, BULB, MB, LOWERED>(grid, stream, cfgs, scenes, nscenes, accum, start,
                                                                  count, counter, slots, event_batch, refill, prof,
                                                                  prop, order, split);
    else
        launch_instance<false, MULTI, LEARNED, BULB, MB, LOWERED>(grid, stream, cfgs, scenes, nscenes, accum, start,
                                                                   count, counter, slots, event_batch, refill, prof,
                                                                   prop, order, split);
}

}  // namespace

hipError_t launch_sdf_persistent(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg, float4 *accum,
                                 uint32_t start, uint32_t count, uint32_t *counter, uint32_t workgroups,
                                 int event_batch, int refill, DeviceProfile *prof, const ProposerDevice *prop,
                                 const uint32_t *tile_order, const SampleSplit &split, hipStream_t stream) {
    uint32_t slots = slot_count(hcfg.width, hcfg.height);
    if (split.groups > 1u) {
        slots *= split.groups;
        const uint32_t npix = hcfg.width * hcfg.height;
        hipLaunchKernelGGL(zero_words_kernel, dim3((npix + 255u) / 256u), dim3(256), 0, stream, split.done, npix);
    }
    if (event_batch <= 0)
        event_batch = prop ? kLearnedEventBatch
                      : hcfg.sdf_id == FPT_SDF_MANDELBULBER
                          ? kMandelbulberEventBatch
                          : default_event_batch(persistent_uses_stash(hcfg, false), split.groups > 1u);
    hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(64), 0, stream, counter);
    const ProposerDevice none{};
    const dim3 grid(workgroups);
    const bool p = prof != nullptr, lowered = has_lowered_tables(hcfg);
#define FPT_LAUNCH(LEARNED, BULB, MB, LOWERED, PROP)                                                              \
    launch_profiled<false, LEARNED, BULB, MB, LOWERED>(p, grid, stream, dcfg, nullptr, 1u, accum, start, count,  \
                                                       counter, slots, event_batch, refill, prof, PROP, tile_order,  \
                                                       split)
    if (prop && lowered) FPT_LAUNCH(true, false, false, true, *prop);
    else if (prop) FPT_LAUNCH(true, false, false, false, *prop);
    else if (hcfg.sdf_id == FPT_SDF_MANDELBULBER) FPT_LAUNCH(false, false, true, false, none);
    // Tables take precedence over the program (SM:3576-3584), so a config
    // that carries them never runs the single-leaf instance.
    else if (lowered) FPT_LAUNCH(false, false, false, true, none);
    else if (single_bulb_program(hcfg)) FPT_LAUNCH(false, true, false, false, none);
    else FPT_LAUNCH(false, false, false, false, none);
#undef FPT_LAUNCH
    return hipGetLastError();
}

hipError_t launch_sdf_batch(const FptRenderConfig *dcfgs, const BatchScene *dscenes, uint32_t nscenes,
                            uint32_t total_slots, float4 *accum, uint32_t *counter, uint32_t workgroups,
                            int event_batch, uint32_t spp_cap, DeviceProfile *prof, bool lowered,
                            const uint32_t *tile_order, hipStream_t stream) {
    hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(64), 0, stream, counter);
    const ProposerDevice none{};
    const bool p = prof != nullptr;
    if (event_batch <= 0) event_batch = default_event_batch(uses_path_stash(true, false, false), true);
    if (lowered)
        launch_profiled<true, false, false, false, true>(p, dim3(workgroups), stream, dcfgs, dscenes, nscenes, accum,
                                                         0u, spp_cap, counter, total_slots, event_batch, kBatchRefill, prof, none,
                                                         tile_order, SampleSplit{});
    else
        launch_profiled<true, false, false, false, false>(p, dim3(workgroups), stream, dcfgs, dscenes, nscenes, accum,
                                                          0u, spp_cap, counter, total_slots, event_batch, kBatchRefill, prof, none,
                                                          tile_order, SampleSplit{});
    return hipGetLastError();
}

}  // namespace fpt
