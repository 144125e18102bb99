// Internal interfaces between the HIP kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

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
};

hipError_t launch_sdf_accumulate(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg, float4 *accum,
                                 uint32_t start, uint32_t count, DeviceProfile *prof, hipStream_t stream);
hipError_t launch_trace_primary(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg, FptRayHit *hits,
                                hipStream_t stream);
hipError_t launch_field_sample(const FptRenderConfig *dcfg, const float4 *pts, size_t n,
                               FptMandelbulberFieldSample *out, hipStream_t stream);
hipError_t launch_surface_sample(const FptRenderConfig *dcfg, const float4 *pts, size_t n, float4 *out,
                                 hipStream_t stream);
hipError_t launch_sdf_persistent(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg, float4 *accum,
                                 uint32_t start, uint32_t count, uint32_t *counter, uint32_t workgroups,
                                 int event_batch, DeviceProfile *prof, hipStream_t stream);
hipError_t launch_present(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg, const float4 *accum,
                          uchar4 *out, hipStream_t stream);

}  // namespace fpt
