// Host launchers of the persistent SDF path tracer (kernel in
// persistent_kernel.h, shared with the runtime-specialised build).
#include <hip/hip_runtime.h>

#include "persistent_kernel.h"

namespace fpt {

hipError_t launch_zero_counter(uint32_t *counter, hipStream_t stream) {
    hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(64), 0, stream, counter);
    return hipGetLastError();
}

uint32_t persistent_slots(uint32_t width, uint32_t height) { return slot_count(width, height); }

bool single_bulb_program(const FptRenderConfig &c) {
    return FPT_INTERLEAVED_BULB && c.sdf_id == FPT_SDF_PROGRAM && c.sdf_program_count == 1u &&
           c.sdf_program[0].opcode == FPT_SDF_OP_MANDELBULB;
}

hipError_t launch_sdf_persistent(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg, float4 *accum,
                                 uint32_t start, uint32_t count, uint32_t *counter, uint32_t workgroups,
                                 int event_batch, DeviceProfile *prof, const ProposerDevice *prop,
                                 hipStream_t stream) {
    const uint32_t slots = slot_count(hcfg.width, hcfg.height);
    hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(64), 0, stream, counter);
    const ProposerDevice none{};
    const dim3 grid(workgroups), block(256);
    if (prop) {
        if (prof)
            hipLaunchKernelGGL((sdf_persistent_kernel<true, false, true>), grid, block, 0, stream, dcfg, nullptr, 1u,
                               accum, start, count, counter, slots, event_batch, prof, *prop);
        else
            hipLaunchKernelGGL((sdf_persistent_kernel<false, false, true>), grid, block, 0, stream, dcfg, nullptr, 1u,
                               accum, start, count, counter, slots, event_batch, nullptr, *prop);
    } else if (hcfg.sdf_id == FPT_SDF_MANDELBULBER) {
        if (prof)
            hipLaunchKernelGGL((sdf_persistent_kernel<true, false, false, false, true>), grid, block, 0, stream, dcfg,
                               nullptr, 1u, accum, start, count, counter, slots, event_batch, prof, none);
        else
            hipLaunchKernelGGL((sdf_persistent_kernel<false, false, false, false, true>), grid, block, 0, stream,
                               dcfg, nullptr, 1u, accum, start, count, counter, slots, event_batch, nullptr, none);
    } else if (single_bulb_program(hcfg)) {
        if (prof)
            hipLaunchKernelGGL((sdf_persistent_kernel<true, false, false, true>), grid, block, 0, stream, dcfg, nullptr,
                               1u, accum, start, count, counter, slots, event_batch, prof, none);
        else
            hipLaunchKernelGGL((sdf_persistent_kernel<false, false, false, true>), grid, block, 0, stream, dcfg,
                               nullptr, 1u, accum, start, count, counter, slots, event_batch, nullptr, none);
    } else if (prof) {
        hipLaunchKernelGGL((sdf_persistent_kernel<true, false, false>), grid, block, 0, stream, dcfg, nullptr, 1u,
                           accum, start, count, counter, slots, event_batch, prof, none);
    } else {
        hipLaunchKernelGGL((sdf_persistent_kernel<false, false, false>), grid, block, 0, stream, dcfg, nullptr, 1u,
                           accum, start, count, counter, slots, event_batch, nullptr, none);
    }
    return hipGetLastError();
}

hipError_t launch_sdf_batch(const FptRenderConfig *dcfgs, const BatchScene *dscenes, uint32_t nscenes,
                            uint32_t total_slots, float4 *accum, uint32_t *counter, uint32_t workgroups,
                            int event_batch, uint32_t spp_cap, DeviceProfile *prof, hipStream_t stream) {
    hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(64), 0, stream, counter);
    const ProposerDevice none{};
    if (prof)
        hipLaunchKernelGGL((sdf_persistent_kernel<true, true, false>), dim3(workgroups), dim3(256), 0, stream, dcfgs,
                           dscenes, nscenes, accum, 0u, spp_cap, counter, total_slots, event_batch, prof, none);
    else
        hipLaunchKernelGGL((sdf_persistent_kernel<false, true, false>), dim3(workgroups), dim3(256), 0, stream, dcfgs,
                           dscenes, nscenes, accum, 0u, spp_cap, counter, total_slots, event_batch, nullptr, none);
    return hipGetLastError();
}

}  // namespace fpt
