// Voxel backend (placeholder until the brick build and DDA land).
#include <cstdio>

#include "voxel_api.h"

namespace fpt {

int voxel_build(const FptRenderConfig &, const FptRenderConfig *, VoxelStore &, hipStream_t, char *error,
                size_t error_len) {
    if (error && error_len) std::snprintf(error, error_len, "%s", "voxel backend is not built yet");
    return 1;
}

void voxel_free(VoxelStore &store) { store = VoxelStore(); }

hipError_t launch_voxel_accumulate(const FptRenderConfig *, const FptRenderConfig &, const VoxelStore &, float4 *,
                                   uint32_t, uint32_t, hipStream_t) {
    return hipErrorNotSupported;
}

}  // namespace fpt

extern "C" int fpt_test_voxel_dda(const char *, char *error, size_t error_len) {
    if (error && error_len) std::snprintf(error, error_len, "%s", "voxel DDA contract is not built yet");
    return 1;
}
