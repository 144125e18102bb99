// Voxel backend interfaces (sparse-brick build + exact grid DDA).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fpt_abi.h"

namespace fpt {

// Device-resident voxel storage of one plan.
struct VoxelStore {
    uint32_t storage = 0;         // FptVoxelStorage
    void *page_table = nullptr;   // u32 per brick: 0 empty, else page + 1
    void *pool = nullptr;         // 12-byte records (4^3 bricks or dense)
    void *distance = nullptr;     // 8^3 layout: FP16 distance per cell
    void *material = nullptr;     // 8^3 layout: u8 material id per cell
    void *palette = nullptr;      // 8^3 layout: 256 x 8 floats
    void *focus = nullptr;        // focus distance of the current render
    uint32_t bricks_per_axis = 0;
    uint32_t brick_size = 0;      // cells per brick edge (4 or 8); 0 dense
    uint32_t active_bricks = 0;
    uint32_t rejected_bricks = 0; // bricks dropped by certified whole-brick rejection
    uint64_t active_cells = 0;
    uint64_t memory_bytes = 0;
};

int voxel_build(const FptRenderConfig &cfg, const FptRenderConfig *dcfg, VoxelStore &store, hipStream_t stream,
                char *error, size_t error_len);
void voxel_free(VoxelStore &store);
hipError_t launch_voxel_accumulate(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg, VoxelStore &store,
                                   float4 *accum, uint32_t start, uint32_t count, hipStream_t stream,
                                   unsigned long long *dda);  // optional {visited, page reads, payload reads}
hipError_t launch_voxel_diagnostic(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg, VoxelStore &store,
                                   const FptDiagnosticConfig &dg, uchar4 *out, hipStream_t stream);
hipError_t launch_voxel_trace_cells(const FptRenderConfig *dcfg, const VoxelStore &store, const float *rays, size_t n,
                                    uint32_t cap, int *cells, uint32_t *counts, float *hits, hipStream_t stream);

}  // namespace fpt
