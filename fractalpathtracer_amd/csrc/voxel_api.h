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
    void *counter = nullptr;      // pixel-slot counter of the persistent voxel kernel
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
// Event batch of the persistent voxel kernel given 0.
constexpr int kVoxelEventBatch = 32;
// dda (optional): {visited, page reads, payload reads}.  persistent: configs
// with stored materials and face normals run the path-regenerating kernel
// (workgroups 0 = the default grid, event_batch 0 = kVoxelEventBatch);
// otherwise, and for configs that evaluate the field, one pixel per lane.
hipError_t launch_voxel_accumulate(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg, VoxelStore &store,
                                   float4 *accum, uint32_t start, uint32_t count, hipStream_t stream,
                                   unsigned long long *dda, bool persistent = true, uint32_t workgroups = 0,
                                   int event_batch = 0);
hipError_t launch_voxel_diagnostic(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg, VoxelStore &store,
                                   const FptDiagnosticConfig &dg, uchar4 *out, hipStream_t stream);
hipError_t launch_voxel_trace_cells(const FptRenderConfig *dcfg, const VoxelStore &store, const float *rays, size_t n,
                                    uint32_t cap, int *cells, uint32_t *counts, float *hits, hipStream_t stream);

}  // namespace fpt
