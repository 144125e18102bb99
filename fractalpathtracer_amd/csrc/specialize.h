// Topology-specialised persistent kernels (host side; specialize.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>

#include "../../include/fpt_abi.h"

namespace fpt {

struct SampleSplit;

struct SpecializedModule {
    hipModule_t module = nullptr;
    hipFunction_t kernel = nullptr;    // sdf_persistent_kernel<false, false, false>
    hipFunction_t validate = nullptr;  // fpt_topology_validate
    hipFunction_t sample = nullptr;    // fpt_topology_sample
    size_t code_bytes = 0;
};

struct SpecializeReport {
    uint32_t cache_status = 0;  // 1 compiled (and written to the cache), 2 loaded from a cache
    uint64_t source_bytes = 0, code_bytes = 0;
    double compile_ms = 0.0, link_ms = 0.0, validate_ms = 0.0;
    uint64_t sample_count = 0, distance_failures = 0, gradient_failures = 0;
    float max_distance_error = 0.f, max_gradient_error = 0.f;
};

// Generated source of `cfg`'s program topology; empty with `failure` set when
// the program cannot be specialised.
std::string topology_source(const FptRenderConfig &cfg, std::string &failure);
// Compile (or fetch from the memory / disk cache) and load the module.
int specialize(const FptRenderConfig &cfg, std::shared_ptr<SpecializedModule> &out, SpecializeReport &rep,
               std::string &failure);
hipError_t validate_specialized(const SpecializedModule &m, const FptRenderConfig *dcfg, uint32_t points,
                                SpecializeReport &rep, hipStream_t stream);
// Specialised distance + gradient at n device points (float4 xyz_) into out.
hipError_t sample_specialized(const SpecializedModule &m, const FptRenderConfig *dcfg, const float4 *pts, uint32_t n,
                              float4 *out, hipStream_t stream);
hipError_t launch_specialized(const SpecializedModule &m, const FptRenderConfig *dcfg, float4 *accum, uint32_t start,
                              uint32_t count, uint32_t *counter, uint32_t slots, uint32_t workgroups, int event_batch,
                              int refill, const uint32_t *tile_order, const SampleSplit &split, hipStream_t stream);

}  // namespace fpt
