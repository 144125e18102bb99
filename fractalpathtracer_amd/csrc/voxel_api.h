// Voxel backend interfaces (sparse-brick build + exact grid DDA).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fpt_abi.h"

namespace fpt {

// Device-resident brick storage of one plan.
struct VoxelStore {
    void *page_table = nullptr;  // u32 per brick: 0 empty, else page + 1
    void *pool = nullptr;        // brick payloads, page-major
    uint32_t bricks_per_axis = 0;
    uint32_t brick_size = 0;     // cells per brick edge (4 or 8)
    uint32_t active_bricks = 0;
    uint64_t active_cells = 0;
    uint64_t memory_bytes = 0;
};

int voxel_build(const FptRenderConfig &cfg, const FptRenderConfig *dcfg, VoxelStore &store, hipStream_t stream,
                char *error, size_t error_len);
void voxel_free(VoxelStore &store);
hipError_t launch_voxel_accumulate(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg, const VoxelStore &store,
                                   float4 *accum, uint32_t start, uint32_t count, hipStream_t stream);

}  // namespace fpt
