/*
 * fpt_abi.h — C ABI of the MI355X (gfx950) hot path of the FPT fractal path
 * tracer.
 *
 * Two groups of declarations live here:
 *
 *   1. The reference's drop-in surface.  Every struct below that carries an
 *      `Fpt` name has the exact field order, field types and size of the
 *      struct of the same name in src/metal_bridge.h of the reference
 *      (metal_bridge.h:58-397), so that src/ffi.rs:1-521 links against
 *      libfpt_hip.so unchanged.  The entry points keep their names and
 *      signatures (metal_bridge.h:404-486).  Arguments that name Metal
 *      artefacts (metallib, stitch library/archive, MSL shader source) are
 *      accepted and ignored: the kernels are compiled for gfx950 ahead of
 *      time.
 *
 *   2. New entry points prefixed `fpt_hip_`.  They have no counterpart in the
 *      reference; each comment says which reference behaviour it exposes
 *      (per-ray float outputs, device-resident rendering, voxel tables, the
 *      batched multi-scene launch, the learned step proposer).
 *
 * Conventions (shared by every entry point, as in metal_bridge.mm:41-47):
 *   - return 0 on success, 1 on failure; on failure a NUL-terminated message
 *     is written to error[0..error_len) when error != NULL;
 *   - the caller owns every pointer; optional out-pointers may be NULL;
 *   - calls are synchronous unless the comment says otherwise.
 */
#ifndef FPT_ABI_H
#define FPT_ABI_H

#ifndef __HIPCC_RTC__
#include <stddef.h>
#include <stdint.h>
#else /* device-only build by the HIP runtime compiler: no C library headers */
typedef unsigned char uint8_t;
typedef unsigned short uint16_t;
typedef unsigned int uint32_t;
typedef unsigned long long uint64_t;
typedef int int32_t;
typedef long long int64_t;
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Scene selector (metal_bridge.h:10-24). */
typedef enum FptSdfId {
    FPT_SDF_CORNELL_BOX = 0,
    FPT_SDF_GLASS_BALL = 1,
    FPT_SDF_BALL_FRACTAL = 2,
    FPT_SDF_CAGE_FRACTAL = 3,
    FPT_SDF_IFS_FRACTAL = 4,
    FPT_SDF_MANDELBOX_FRACTAL = 5,
    FPT_SDF_MENGER_SPONGE = 6,
    FPT_SDF_TOWER_FRACTAL = 7,
    FPT_SDF_TREE_FRACTAL = 8,
    FPT_SDF_PROGRAM = 9,
    FPT_SDF_README_CORNELL = 10,
    FPT_SDF_README_GLASS = 11,
    FPT_SDF_MANDELBULBER = 12
} FptSdfId;

/* Capacities (metal_bridge.h:26-33). */
enum {
    FPT_SDF_PROGRAM_MAX_OPS = 64,
    FPT_SDF_FLAT_UNION_MAX_PRIMITIVES = 64,
    FPT_SDF_GRADIENT_MAX_STOPS = 16,
    FPT_HDRI_PATH_CAPACITY = 256,
    FPT_HDRI_LUT_WIDTH = 32,
    FPT_HDRI_LUT_HEIGHT = 16
};

enum {
    FPT_SDF_STITCH_STATE_FULL = 0,
    FPT_SDF_STITCH_STATE_LEAN = 1
};

/*
 * Typed SDF program opcodes (metal_bridge.h:40-56).  Values 1-8, 16-18, 24-26
 * and 32 are the reference's.  The fractal opcodes below use numbers the
 * reference leaves unassigned; the reference has no fractal leaves in its
 * program vocabulary (its fractals are fixed sdf_id scenes).
 *
 *   FPT_SDF_OP_BOX_FOLD   (transform)  p = clamp(p, -l, l) * 2 - p   with
 *                         l = data[0] (0 -> 1).  The Mandelbulber box fold
 *                         (fold limit l, fold value 2l), applied once to the
 *                         running point, Lipschitz-1.
 *   FPT_SDF_OP_MANDELBULB (leaf)  data = {power, iterations, bailout, -}.
 *                         Triplex power z <- r^p (cos t cos f, cos t sin f,
 *                         sin t) + c with t = asin(z/r), f = atan2(y, x);
 *                         dr <- p r^(p-1) dr + 1; DE = 0.5 r ln r / dr
 *                         (src/mandelbulber/orbit.rs:766-778, 530-585,
 *                         654-660).
 *   FPT_SDF_OP_MANDELBOX  (leaf)  data = {scale, iterations, min_radius,
 *                         fixed_radius}.  Box fold (limit 1, value 2),
 *                         sphere fold on squared radii, z <- scale z + c,
 *                         dr <- |scale| dr + 1, bailout 100, DE = r / dr
 *                         (orbit.rs:1636-1687 without rotations,
 *                         src/mandelbulber.rs:1583-1640).
 *   FPT_SDF_OP_MENGER     (leaf)  data = {iterations, size, divisor, -}.
 *                         Cube of half size `size` minus `iterations` levels
 *                         of axis crosses whose period shrinks by `divisor`
 *                         (defaults 1 and 3) (Shaders.metal:756-784 with a
 *                         parametric iteration count and no inter-level
 *                         rotation).
 *
 * Leaf `flags` select the CSG combine exactly like sphere/box/plane:
 * 0 union (min), 1 intersection (max), 2 subtraction (max(a, -b)).
 */
typedef enum FptSdfOpcode {
    FPT_SDF_OP_ABS = 1,
    FPT_SDF_OP_TRANSLATE = 2,
    FPT_SDF_OP_SCALE = 3,
    FPT_SDF_OP_ROTATE_X = 4,
    FPT_SDF_OP_ROTATE_Y = 5,
    FPT_SDF_OP_ROTATE_Z = 6,
    FPT_SDF_OP_REPEAT = 7,
    FPT_SDF_OP_SORT_DESC = 8,
    FPT_SDF_OP_BOX_FOLD = 9,
    FPT_SDF_OP_SPHERE = 16,
    FPT_SDF_OP_BOX = 17,
    FPT_SDF_OP_PLANE = 18,
    FPT_SDF_OP_MANDELBULB = 19,
    FPT_SDF_OP_MANDELBOX = 20,
    FPT_SDF_OP_MENGER = 21,
    FPT_SDF_OP_UNION = 24,
    FPT_SDF_OP_INTERSECTION = 25,
    FPT_SDF_OP_SUBTRACT = 26,
    FPT_SDF_OP_ORBIT_ADD = 32
} FptSdfOpcode;

/* One program instruction, 32 bytes (metal_bridge.h:58-64). */
struct FptSdfInstruction {
    uint32_t opcode;
    uint32_t flags;
    uint32_t material_index;
    uint32_t _pad0;
    float data[4];
};

/* metal_bridge.h:66-73 (80 bytes). */
struct FptPrimitiveInstance {
    float transform[12];
    float data[4];
    uint32_t opcode;
    float distance_scale;
    uint32_t source_instruction;
    uint32_t _pad0;
};

/* metal_bridge.h:75-79 (64 bytes). */
struct FptAffineTransform {
    float transform[12];
    float distance_scale;
    uint32_t _pad0[3];
};

/* metal_bridge.h:81-87 (32 bytes). */
struct FptIndexedPrimitive {
    float data[4];
    uint32_t opcode;
    uint32_t transform_index;
    uint32_t source_instruction;
    uint32_t combine_mode;
};

/* metal_bridge.h:89-114 (5136 bytes). */
struct FptTypedSoAProgram {
    uint32_t sphere_count;
    uint32_t box_count;
    uint32_t plane_count;
    uint32_t _pad0;
    float sphere_x[FPT_SDF_FLAT_UNION_MAX_PRIMITIVES];
    float sphere_y[FPT_SDF_FLAT_UNION_MAX_PRIMITIVES];
    float sphere_z[FPT_SDF_FLAT_UNION_MAX_PRIMITIVES];
    float sphere_radius[FPT_SDF_FLAT_UNION_MAX_PRIMITIVES];
    uint32_t sphere_source[FPT_SDF_FLAT_UNION_MAX_PRIMITIVES];
    float box_x[FPT_SDF_FLAT_UNION_MAX_PRIMITIVES];
    float box_y[FPT_SDF_FLAT_UNION_MAX_PRIMITIVES];
    float box_z[FPT_SDF_FLAT_UNION_MAX_PRIMITIVES];
    float box_half_x[FPT_SDF_FLAT_UNION_MAX_PRIMITIVES];
    float box_half_y[FPT_SDF_FLAT_UNION_MAX_PRIMITIVES];
    float box_half_z[FPT_SDF_FLAT_UNION_MAX_PRIMITIVES];
    uint32_t box_source[FPT_SDF_FLAT_UNION_MAX_PRIMITIVES];
    float plane_x[FPT_SDF_FLAT_UNION_MAX_PRIMITIVES];
    float plane_y[FPT_SDF_FLAT_UNION_MAX_PRIMITIVES];
    float plane_z[FPT_SDF_FLAT_UNION_MAX_PRIMITIVES];
    float plane_center_x[FPT_SDF_FLAT_UNION_MAX_PRIMITIVES];
    float plane_center_y[FPT_SDF_FLAT_UNION_MAX_PRIMITIVES];
    float plane_center_z[FPT_SDF_FLAT_UNION_MAX_PRIMITIVES];
    float plane_offset[FPT_SDF_FLAT_UNION_MAX_PRIMITIVES];
    uint32_t plane_source[FPT_SDF_FLAT_UNION_MAX_PRIMITIVES];
};

typedef enum FptSdfAccumulationMode {
    FPT_SDF_ACCUMULATION_AUTO = 0,
    FPT_SDF_ACCUMULATION_PER_SAMPLE = 1,
    FPT_SDF_ACCUMULATION_BATCH = 2,
    FPT_SDF_ACCUMULATION_CHUNKED = 3
} FptSdfAccumulationMode;

typedef enum FptRendererBackend {
    FPT_RENDERER_SDF = 0,
    FPT_RENDERER_VOXEL = 1,
    FPT_RENDERER_BOUND_GRID = 2,
    FPT_RENDERER_REGIONAL = 3
} FptRendererBackend;

typedef enum FptVoxelNormalMode {
    FPT_VOXEL_NORMAL_FACE = 0,
    FPT_VOXEL_NORMAL_SMOOTH = 1,
    FPT_VOXEL_NORMAL_EXACT = 2
} FptVoxelNormalMode;

typedef enum FptVoxelMaterialMode {
    FPT_VOXEL_MATERIAL_STORED = 0,
    FPT_VOXEL_MATERIAL_EXACT = 1
} FptVoxelMaterialMode;

typedef enum FptVoxelOffsetMode {
    FPT_VOXEL_OFFSET_LEGACY = 0,
    FPT_VOXEL_OFFSET_PRECISION = 1
} FptVoxelOffsetMode;

/*
 * FPT_VOXEL_STORAGE_SPARSE_BRICKS8 is new: 8^3 bricks whose cells hold an
 * FP16 signed distance and a u8 material id (the reference's bricks are 4^3
 * cells of a 12-byte packed colour/material record).
 */
typedef enum FptVoxelStorage {
    FPT_VOXEL_STORAGE_DENSE = 0,
    FPT_VOXEL_STORAGE_SPARSE_BRICKS = 1,
    FPT_VOXEL_STORAGE_TEMPLATE_BRICKS = 3,
    FPT_VOXEL_STORAGE_SPARSE_BRICKS8 = 8
} FptVoxelStorage;

/*
 * FPT_VOXEL_COVERAGE_SHELL is new: legacy occupancy (|d| <= r * band), then
 * every occupied cell whose 26 neighbours are all occupied is dropped.  The
 * first occupied cell a DDA meets from an unoccupied cell is always kept (its
 * predecessor is one of its 26 neighbours), so hits of rays that start
 * outside the solid are unchanged while a solid fractal interior costs no
 * storage.
 */
typedef enum FptVoxelCoverageMode {
    FPT_VOXEL_COVERAGE_LEGACY = 0,
    FPT_VOXEL_COVERAGE_LIPSCHITZ = 1,
    FPT_VOXEL_COVERAGE_INTERVAL = 2,
    FPT_VOXEL_COVERAGE_SHELL = 3
} FptVoxelCoverageMode;

typedef enum FptVoxelBuildMode {
    FPT_VOXEL_BUILD_STAGING = 0,
    FPT_VOXEL_BUILD_DIRECT = 1
} FptVoxelBuildMode;

typedef enum FptVoxelLeafRefinement {
    FPT_VOXEL_LEAF_REFINEMENT_NONE = 0,
    FPT_VOXEL_LEAF_REFINEMENT_SECANT_BISECTION = 1,
    FPT_VOXEL_LEAF_REFINEMENT_RESTRICTED_TRACE = 2,
    FPT_VOXEL_LEAF_REFINEMENT_FIXED_DE = 3
} FptVoxelLeafRefinement;

/*
 * The whole render description, 30,448 bytes (metal_bridge.h:170-267).
 * Field meanings follow the reference (src/scene.rs:1651-1724):
 *   render[8] = {bounces, marching_steps, normal_epsilon, min_distance,
 *                max_distance, adaptive_marching, error_protection,
 *                random_mode}
 *   world[7]  = {mode, light_size, rotation, elevation, power, contrast,
 *                background_mask}
 *   sun[5]    = {enabled, rotation_deg, elevation_deg, power, softness}
 */
struct FptRenderConfig {
    uint32_t width;
    uint32_t height;
    uint32_t samples;
    uint32_t preview;
    uint32_t sdf_id;
    uint32_t glass_mode;
    uint32_t sdf_accumulation_mode;
    uint32_t sdf_profile;
    uint32_t sdf_bounce_cap;
    uint32_t sdf_russian_roulette;
    uint32_t sdf_normal_mode;
    uint32_t sdf_chunk_samples;
    float sdf_rr_start;
    float sdf_rr_min_prob;
    float mandel_iteration_scale;

    float camera_position[3];
    float camera_yaw_pitch[2];
    float camera_roll;
    float camera_fov;
    float camera_dof;
    float focus_distance;

    float render[8];
    float world[7];
    float world_one_color[3];
    float sun[5];
    float sun_color[3];
    float background_gradient[6];
    float post[7];
    float set_values[40];
    float vset_values[133];

    uint32_t sdf_program_count;
    uint32_t gradient_count;
    uint32_t material_mode;
    uint32_t hdri_enabled;
    uint32_t hdri_width;
    uint32_t hdri_height;
    uint32_t _pad_hdri[2];
    uint32_t fractal_style_mode;
    uint32_t _pad_style[3];
    float fractal_style[12];
    float program_material[8];
    struct FptSdfInstruction sdf_program[FPT_SDF_PROGRAM_MAX_OPS];
    float gradient_stops[FPT_SDF_GRADIENT_MAX_STOPS][4];
    uint8_t hdri_path[FPT_HDRI_PATH_CAPACITY];
    uint16_t hdri_lut[FPT_HDRI_LUT_WIDTH * FPT_HDRI_LUT_HEIGHT * 3];
    uint32_t renderer_backend;
    uint32_t voxel_resolution;
    uint32_t voxel_normal_mode;
    uint32_t voxel_storage;
    float voxel_bounds_min[3];
    float voxel_surface_band;
    float voxel_bounds_max[3];
    uint32_t voxel_fill_interior;
    uint32_t voxel_coverage_mode;
    uint32_t voxel_build_mode;
    uint32_t voxel_brick_rejection;
    uint32_t voxel_leaf_refinement;
    uint32_t voxel_material_mode;
    uint32_t voxel_offset_mode;
    uint32_t bound_grid_resolution;
    uint32_t bound_grid_profile;
    uint32_t bound_grid_profile_stride;
    uint32_t bound_grid_cage_bounds;
    uint32_t bound_grid_directional;
    uint32_t sdf_program_source_count;
    uint32_t sdf_program_optimization;
    uint32_t regional_program_resolution;
    uint32_t bound_grid_fp16;
    uint32_t sdf_flat_union_count;
    struct FptPrimitiveInstance
        sdf_flat_union_instances[FPT_SDF_FLAT_UNION_MAX_PRIMITIVES];
    uint32_t sdf_shading_program_count;
    uint32_t sdf_geometry_split;
    struct FptSdfInstruction sdf_shading_program[FPT_SDF_PROGRAM_MAX_OPS];
    uint32_t sdf_topology_specialization;
    uint32_t sdf_runtime_source_bytecode;
    uint32_t sdf_function_stitching;
    uint32_t sdf_stitched_surface;
    uint32_t sdf_stitch_validation;
    uint32_t sdf_stitch_distance_only;
    uint32_t sdf_stitch_split_graph;
    uint32_t sdf_stitch_fusion;
    uint32_t sdf_canonical_count;
    uint32_t sdf_canonical_source_count;
    struct FptPrimitiveInstance
        sdf_canonical_primitives[FPT_SDF_FLAT_UNION_MAX_PRIMITIVES];
    uint32_t sdf_canonical_transform_count;
    uint32_t _pad_canonical_indexed[3];
    struct FptAffineTransform
        sdf_canonical_transforms[FPT_SDF_FLAT_UNION_MAX_PRIMITIVES];
    struct FptIndexedPrimitive
        sdf_indexed_primitives[FPT_SDF_FLAT_UNION_MAX_PRIMITIVES];
    struct FptTypedSoAProgram sdf_typed_soa;
};

/* metal_bridge.h:269-309.  Accepted, zero-filled (bound-grid backend is
 * not part of this library). */
struct FptBoundGridStats {
    uint64_t macro_cells[3];
    uint64_t certified_skips[3];
    uint64_t candidate_intervals[3];
    uint64_t candidate_misses[3];
    uint64_t candidate_hits[3];
    uint64_t unknown_intervals[3];
    uint64_t field_evaluations[3];
    uint64_t directional_steps[3];
    uint64_t cell_exit_clamps[3];
    uint64_t unknown_derivative_intervals[3];
    uint64_t profiled_paths;
    uint64_t certified_cells;
    uint64_t unknown_cells;
    uint64_t sampled_bound_failures;
    uint64_t sampled_false_skips;
    uint64_t certified_derivative_cells;
    uint64_t unknown_derivative_cells;
    uint64_t sampled_derivative_failures;
    uint64_t regional_cells;
    uint64_t regional_fallback_cells;
    uint64_t regional_unique_programs;
    uint64_t regional_retained_instructions;
    uint64_t regional_sampled_distance_failures;
    uint64_t regional_profiled_paths;
    uint64_t regional_distance_evaluations[3];
    uint64_t regional_atlas_evaluations[3];
    uint64_t regional_full_program_evaluations[3];
    uint64_t regional_cell_entries[3];
    uint64_t regional_same_cell_reuses[3];
    uint64_t regional_same_program_reuses[3];
    uint64_t regional_program_id_loads[3];
    uint64_t regional_header_loads[3];
    uint64_t regional_dynamic_instructions[3];
    uint64_t regional_pruned_cells;
    uint64_t regional_pruned_primitives;
    uint64_t regional_repeat_seam_fallback_cells;
    uint64_t regional_unsupported_fallback_cells;
    uint64_t regional_no_dominance_cells;
};

/* metal_bridge.h:311-317. */
struct FptStitchValidationStats {
    uint64_t sample_count;
    uint64_t distance_failures;
    uint64_t gradient_failures;
    float max_distance_error;
    float max_gradient_error;
};

/* metal_bridge.h:319-334 (72 bytes). */
struct FptStitchPipelineStats {
    uint32_t instruction_count;
    uint32_t transform_instruction_count;
    uint32_t primitive_instruction_count;
    uint32_t primitive_type_runs;
    uint32_t union_count;
    uint32_t intersection_count;
    uint32_t subtraction_count;
    uint32_t thread_execution_width;
    uint32_t max_total_threads_per_threadgroup;
    uint32_t graph_node_count;
    uint64_t static_threadgroup_memory_bytes;
    uint64_t runtime_source_bytes;
    double runtime_library_compile_ms;
    double runtime_pipeline_link_ms;
};

/* Work counters of one render (metal_bridge.h:336-358, 288 bytes). */
struct FptSdfProfileStats {
    uint64_t primary_steps;
    uint64_t secondary_steps;
    uint64_t shadow_steps;
    uint64_t normal_evals;
    uint64_t bounces;
    uint64_t pixels;
    uint64_t distance_evals;
    uint64_t march_orbit_iterations;
    uint64_t refinement_steps;
    uint64_t normal_field_evals;
    uint64_t material_evals;
    uint64_t max_ray_steps;
    uint64_t max_pixel_steps;
    uint64_t distance_evals_by_phase[4];
    uint64_t orbit_iterations_by_phase[4];
    uint64_t formula_slot_iterations[9];
    uint64_t refinement_distance_evals;
    double primary_ms_estimate;
    double secondary_ms_estimate;
    double shadow_ms_estimate;
    double normal_ms_estimate;
    double bounce_ms_estimate;
};

/* metal_bridge.h:360-368 (40 bytes). */
struct FptAsyncJitStats {
    double fallback_render_ms;
    double jit_build_ms;
    double stitched_render_ms;
    uint32_t cache_status;
    uint32_t fallback_completed_before_jit;
    float max_absolute_error;
    uint32_t _pad0;
};

typedef enum FptDiagnosticMode {
    FPT_DIAGNOSTIC_DEPTH = 0,
    FPT_DIAGNOSTIC_NORMAL = 1,
    FPT_DIAGNOSTIC_MATERIAL = 2,
    FPT_DIAGNOSTIC_HIT_MASK = 3,
    FPT_DIAGNOSTIC_PATH_DIRECT = 4,
    FPT_DIAGNOSTIC_PATH_ENVIRONMENT = 5,
    FPT_DIAGNOSTIC_PATH_THROUGHPUT = 6,
    FPT_DIAGNOSTIC_PATH_FINAL = 7,
    FPT_DIAGNOSTIC_DIFFUSE_NORMAL = 8,
    FPT_DIAGNOSTIC_MANDEL_COLOR_INDEX = 9,
    FPT_DIAGNOSTIC_MANDEL_PALETTE_POSITION = 10,
    FPT_DIAGNOSTIC_SDF_PRIMARY_STEPS = 11,
    FPT_DIAGNOSTIC_SDF_SHADOW_STEPS = 12,
    FPT_DIAGNOSTIC_SDF_NORMAL_EVALS = 13,
    FPT_DIAGNOSTIC_SDF_BOUNCES = 14,
    FPT_DIAGNOSTIC_SDF_BOUNCE_CONTRIBUTION = 15
} FptDiagnosticMode;

/* metal_bridge.h:389-395 (24 bytes). */
struct FptDiagnosticConfig {
    uint32_t mode;
    uint32_t _pad0;
    float max_distance;
    float normal_mix;
    uint32_t dispatch_origin[2];
};

/* metal_bridge.h:397-402 (16 bytes). */
struct FptMandelbulberFieldSample {
    float distance;
    float radius;
    float derivative;
    float iterations;
};

#ifdef __cplusplus
#define FPT_ABI_ASSERT(cond, msg) static_assert(cond, msg)
#else
#define FPT_ABI_ASSERT(cond, msg) _Static_assert(cond, msg)
#endif
/* Sizes pinned by src/ffi.rs:494-507 and metal_bridge.mm:32-39. */
FPT_ABI_ASSERT(sizeof(struct FptSdfInstruction) == 32, "FptSdfInstruction");
FPT_ABI_ASSERT(sizeof(struct FptPrimitiveInstance) == 80, "FptPrimitiveInstance");
FPT_ABI_ASSERT(sizeof(struct FptAffineTransform) == 64, "FptAffineTransform");
FPT_ABI_ASSERT(sizeof(struct FptIndexedPrimitive) == 32, "FptIndexedPrimitive");
FPT_ABI_ASSERT(sizeof(struct FptTypedSoAProgram) == 5136, "FptTypedSoAProgram");
FPT_ABI_ASSERT(sizeof(struct FptStitchPipelineStats) == 72, "FptStitchPipelineStats");
FPT_ABI_ASSERT(sizeof(struct FptRenderConfig) == 30448, "FptRenderConfig");
FPT_ABI_ASSERT(sizeof(struct FptSdfProfileStats) == 288, "FptSdfProfileStats");
FPT_ABI_ASSERT(sizeof(struct FptDiagnosticConfig) == 24, "FptDiagnosticConfig");
FPT_ABI_ASSERT(sizeof(struct FptMandelbulberFieldSample) == 16, "FptMandelbulberFieldSample");
FPT_ABI_ASSERT(sizeof(struct FptAsyncJitStats) == 40, "FptAsyncJitStats");

/* ------------------------------------------------------------------------
 * 1. Reference entry points (metal_bridge.h:404-486).
 * ------------------------------------------------------------------------ */

/*
 * Render `config` with the SDF or voxel backend
 * (metal_bridge.mm:4298-5863).  linear_output, when non-NULL, receives the
 * float4 running-mean accumulation buffer, row-major with row 0 at the top
 * of the camera's view (y grows downward in uv), and must hold at least
 * width*height*4 floats.  output_path, when non-NULL and non-empty, receives
 * a PNG of the presented image (vertically flipped as the reference does).
 * elapsed_ms covers the render kernels only; build_ms the voxel build.
 */
int fpt_metal_render(const char *metallib_path,
                     const char *stitch_metallib_path,
                     const char *stitch_archive_path,
                     const char *output_path,
                     const char *shader_source,
                     size_t shader_source_len,
                     const struct FptRenderConfig *config,
                     double *build_ms,
                     double *elapsed_ms,
                     uint64_t *voxel_memory_bytes,
                     uint32_t *voxel_active_bricks,
                     uint64_t *voxel_active_cells,
                     uint32_t *voxel_rejected_bricks,
                     struct FptBoundGridStats *bound_grid_stats,
                     uint32_t *stitch_cache_status,
                     struct FptStitchValidationStats *stitch_validation_stats,
                     struct FptStitchPipelineStats *stitch_pipeline_stats,
                     struct FptSdfProfileStats *sdf_profile_stats,
                     float *linear_output,
                     size_t linear_output_len,
                     char *error,
                     size_t error_len);

/* Per-pixel diagnostic image (metal_bridge.mm:5865-6024). */
int fpt_metal_diagnostic_render(const char *metallib_path,
                                const char *output_path,
                                const char *shader_source,
                                size_t shader_source_len,
                                const struct FptRenderConfig *config,
                                const struct FptDiagnosticConfig *diagnostic,
                                double *elapsed_ms,
                                char *error,
                                size_t error_len);

/*
 * Distance-field samples at points_xyzw[4*i .. 4*i+2] (metal_bridge.mm:
 * 4205-4296).  This library evaluates the scene selected by `config`
 * (typed program or built-in fractal) and reports {distance, radius,
 * derivative, iterations} of the first fractal leaf it meets (radius,
 * derivative and iterations are 0 for scenes without one).
 */
int fpt_mandelbulber_sample_field(const char *metallib_path,
                                  const char *shader_source,
                                  size_t shader_source_len,
                                  const struct FptRenderConfig *config,
                                  const float *points_xyzw,
                                  size_t point_count,
                                  struct FptMandelbulberFieldSample *samples,
                                  char *error,
                                  size_t error_len);

/* Image comparison report (metal_bridge.mm:6062-6145); PNG in, JSON out. */
int fpt_compare_images(const char *baseline_path,
                       const char *candidate_path,
                       const char *report_path,
                       char *error,
                       size_t error_len);

/* Name of the GPU (metal_bridge.h:467). */
int fpt_metal_device_name(char *name, size_t name_len);

/* Voxel DDA known-answer contract (metal_bridge.mm:3715-3926). */
int fpt_test_voxel_dda(const char *metallib_path, char *error, size_t error_len);

/* Typed structure-of-arrays validator (metal_bridge.mm:4116-4203,
 * Shaders.metal:5106-5154): the config's typed table against its
 * interpreted program at 2^20 hashed points of the voxel bounds; stats are
 * filled, and 1 is returned with the failure counts when any point misses
 * 2e-5 (1 + |d|) in distance or 1e-4 in gradient.  metallib_path is
 * ignored. */
int fpt_test_typed_soa(const char *metallib_path,
                       const struct FptRenderConfig *config,
                       struct FptStitchValidationStats *stats,
                       char *error,
                       size_t error_len);

/* The two entry points below exist for link compatibility; the Metal
 * features behind them (function stitching, AppKit preview) are not part
 * of this library and they return 1 with a message. */
int fpt_test_async_stitch_context(const char *metallib_path,
                                  const char *stitch_metallib_path,
                                  const char *stitch_archive_path,
                                  const struct FptRenderConfig *config,
                                  struct FptAsyncJitStats *stats,
                                  char *error,
                                  size_t error_len);

int fpt_metal_preview(const char *metallib_path,
                      const char *stitch_metallib_path,
                      const char *const *stitch_archive_paths,
                      const char *shader_source,
                      size_t shader_source_len,
                      const struct FptRenderConfig *config,
                      const struct FptRenderConfig *scene_configs,
                      uint32_t scene_config_count,
                      uint32_t scene_config_index,
                      char *error,
                      size_t error_len);

/* ------------------------------------------------------------------------
 * 2. New entry points (no counterpart in the reference).
 * ------------------------------------------------------------------------ */

/*
 * Per-ray primary hit record.  The reference only exposes 8-bit encoded
 * depth/normal images (Shaders.metal:7282-7388); parity needs floats.
 */
struct FptRayHit {
    float t;            /* |hit point - camera| */
    float position[3];  /* march end point */
    float normal[3];    /* normalAt(position); zero when hit == 0 */
    uint32_t hit;       /* 1 when the march ended inside 0.99 * max_distance */
    uint32_t steps;     /* march iterations taken */
    uint32_t _pad0;
};
FPT_ABI_ASSERT(sizeof(struct FptRayHit) == 40, "FptRayHit");

/*
 * March the un-jittered primary ray of every pixel with the renderer's own
 * march (Shaders.metal:3753-3783) and report hit t, position and normal
 * (Shaders.metal:3871-3971).  hits[y*width + x], hit_count >= width*height.
 */
int fpt_hip_trace_primary(const struct FptRenderConfig *config,
                          struct FptRayHit *hits,
                          size_t hit_count,
                          char *error,
                          size_t error_len);

/*
 * Program surface of the typed program at points_xyzw[4*i..4*i+2]:
 * out_dgrad[4*i..4*i+3] = {distance, d/dx, d/dy, d/dz}, the analytic
 * gradient normalAt() normalises (Shaders.metal:3266-3369, extended through
 * the fractal leaves by forward-mode differentiation).
 */
int fpt_hip_program_surface(const struct FptRenderConfig *config,
                            const float *points_xyzw,
                            size_t point_count,
                            float *out_dgrad,
                            char *error,
                            size_t error_len);

/*
 * HDRI hydration, in place, of a host config: clears hdri_width, hdri_height
 * and hdri_lut, then, when world[0] == 2 and hdri_enabled, reads the Radiance
 * RGBE file named by hdri_path and box-downsamples it into the 32 x 16 FP16
 * LUT.  Every render entry point (and plan or batch creation) does this on its
 * own copy of the config, as the reference's render calls do
 * (metal_bridge.mm:2262-2300, 4330-4332, 5879-5881); this entry point exposes
 * the same step to callers that want the table itself.  Host only.
 */
int fpt_hip_hydrate_hdri(struct FptRenderConfig *config, char *error, size_t error_len);

/*
 * Scene front end (host only): JSON scene files and .fpt presets into a
 * config, and the typed-program passes run on it before rendering.  These
 * replace the Rust loader src/scene.rs:743-1976 (load_scene_config and the
 * functions it calls), which the reference keeps above its C ABI.
 *
 * FptSceneOptions carries the render arguments that shape a loaded config
 * (RenderArgs, src/scene.rs:199-273); zero-initialised means "as the scene
 * says, no optional passes, SDF renderer".
 */
struct FptSceneOptions {
    uint32_t width, height, samples; /* 0: the scene's value (defaults 768, 768, 64) */
    uint32_t preview, sdf_profile, glass_mode;
    uint32_t program_optimization;   /* 0 off, 1 basic (optimize_sdf_program) */
    uint32_t geometry_split;         /* split_sdf_geometry_program(enabled) */
    uint32_t canonical_ir;           /* compile_canonical_geometry_program */
    uint32_t flat_union;             /* lower_flat_union_program */
    uint32_t typed_soa;              /* 0 off, 1 required, 2 when possible (automatic backend probing) */
    uint32_t renderer_backend;       /* FPT_RENDERER_* of the render that follows */
    uint32_t topology_specialization;
    uint32_t _pad0;
};
FPT_ABI_ASSERT(sizeof(struct FptSceneOptions) == 56, "FptSceneOptions");

/* default_config (src/scene.rs:743-778). */
int fpt_scene_default_config(struct FptRenderConfig *config);
/* load_scene_config (src/scene.rs:1619-1976); output_name receives the
 * scene's "output" or "<preset stem>.png". */
int fpt_scene_load(const char *scene_path, const char *fpt_root, const struct FptSceneOptions *options,
                   struct FptRenderConfig *config, char *output_name, size_t output_name_len, char *error,
                   size_t error_len);
/* load_fpt_settings (src/scene.rs:1583-1617): SET and vector values, and the
 * Gradient(Orbit_Trap) example program. */
int fpt_scene_load_fpt_settings(const char *preset_path, struct FptRenderConfig *config, char *error,
                                size_t error_len);
/* apply_sdf_program (src/scene.rs:963-1021) of one sdf_program JSON object. */
int fpt_scene_apply_program_json(struct FptRenderConfig *config, const char *json, char *error, size_t error_len);
/* optimize_sdf_program (src/scene.rs:1124-1157). */
void fpt_scene_optimize_program(struct FptRenderConfig *config);
/* split_sdf_geometry_program (src/scene.rs:1159-1184). */
void fpt_scene_split_program(struct FptRenderConfig *config, uint32_t enabled);
/* compile_canonical_geometry_program (src/scene.rs:1211-1358); 1 when compiled. */
int fpt_scene_compile_canonical(struct FptRenderConfig *config);
/* lower_flat_union_program (src/scene.rs:1360-1446); 1 when lowered. */
int fpt_scene_lower_flat_union(struct FptRenderConfig *config);
/* lower_typed_soa_program (src/scene.rs:1448-1530); 1 when lowered. */
int fpt_scene_lower_typed_soa(struct FptRenderConfig *config);

/*
 * Device-resident render plan: the config (and any derived device tables:
 * voxel pool, step-proposer weights) uploaded once, so repeated renders
 * touch only HBM.  fpt_hip_plan_render enqueues the render on `stream`
 * (a hipStream_t, NULL for the default stream) into device_accum
 * (width*height float4, device memory) and returns without synchronising.
 */
typedef struct FptHipPlan FptHipPlan;

int fpt_hip_plan_create(const struct FptRenderConfig *config,
                        FptHipPlan **plan,
                        char *error,
                        size_t error_len);
int fpt_hip_plan_render(FptHipPlan *plan,
                        float *device_accum,
                        void *stream,
                        char *error,
                        size_t error_len);
void fpt_hip_plan_destroy(FptHipPlan *plan);

/*
 * Launch shape of a plan's SDF render (the autotuner's knobs; every shape
 * renders the same image):
 *   kernel       0 = one path per lane (16x16 pixel tiles per workgroup);
 *                1 = persistent, path-regenerating kernel (default);
 *   workgroups   persistent grid size in 256-thread workgroups, 0 = two per
 *                compute unit;
 *   event_batch  how many paths of a 64-lane wave must have finished their
 *                march before the per-hit work (material, analytic normal,
 *                sun and scatter) runs for all of them, 1..128.  The
 *                exact-march instances with parked paths (see below; also
 *                batches of scenes) count the wave's lanes and its paths
 *                parked in LDS, up to 128, and also run the per-hit work
 *                once as many paths wait as march; the others count lanes
 *                (values above 64 then wait for every march to end).  0 =
 *                the instance's default: with parked paths 64 for a launch
 *                in sample runs (below) and 48 otherwise, 24 without, 40
 *                for imported Mandelbulber scenes, 16 for the learned march
 *                (fpt_hip_batch_tune: 64 and 24).
 */
int fpt_hip_plan_tune(FptHipPlan *plan,
                      uint32_t kernel,
                      uint32_t workgroups,
                      uint32_t event_batch,
                      char *error,
                      size_t error_len);

/*
 * Path stash and sample runs of the persistent kernel's single-scene
 * exact-march instances (no reference counterpart: launch-shape knobs of
 * this library; images do not depend on them).
 *   refill         each wave parks up to 64 further paths in LDS; lanes whose
 *                  path waits for the per-hit work, or that are idle, exchange
 *                  it for a parked marching path once `refill` lanes of the
 *                  wave are not marching (1..64, default 6);
 *   sample_groups  each pixel's samples of a launch are split into at most
 *                  this many runs of at least 8 samples (1..64, default 8),
 *                  separate work slots that different lanes render at once;
 *                  the run that finishes last folds the per-sample colours
 *                  in sample order.  1 = one slot per pixel.  Applies to
 *                  frames whose width and height are multiples of 8.  The
 *                  plan keeps the per-sample colours on the device (12 B per
 *                  pixel and sample; not used beyond 8 GB).
 */
int fpt_hip_plan_tune_stash(FptHipPlan *plan, uint32_t refill, uint32_t sample_groups, char *error,
                            size_t error_len);

/*
 * Voxel tables of `config` (renderer_backend ignored; voxel_* fields used),
 * built on the GPU exactly as fpt_metal_render builds them
 * (Shaders.metal:5388-5424, metal_bridge.mm:1718-1816; pages are numbered in
 * linear brick order bx + G*(by + G*bz), the reference's staging order).
 * With voxel_build_mode DIRECT and voxel_brick_rejection set, bricks whose
 * certified program bounds exclude the surface keep no page, as the
 * reference's direct build drops them (Shaders.metal:5704-5726); the tables
 * then equal the reference's cell for cell, with pages in staging order
 * where the reference's direct build numbers them in completion order.
 * info is always filled.  Each output buffer is optional (NULL skips it); a
 * buffer that is too small fails the call and the message names the size.
 *   page_table  u32[G^3]                      (sparse layouts)
 *   records     12-byte records (4^3 bricks: pages*64; dense: res^3)
 *   distance    FP16 bits u16[pages*512]       (8^3 layout)
 *   material    u8[pages*512], 0 = empty       (8^3 layout)
 *   palette     float[256*8] {r,g,b,roughness,specular,translucency,ior,
 *               emission} per material id     (8^3 layout)
 */
struct FptVoxelBuildInfo {
    uint32_t storage;
    uint32_t brick_size;        /* 4 or 8; 0 for dense */
    uint32_t bricks_per_axis;   /* G */
    uint32_t active_bricks;     /* pages */
    uint64_t active_cells;
    uint64_t memory_bytes;
    double build_ms;
    uint32_t rejected_bricks;   /* direct build with brick rejection (SM:5704-5726) */
    uint32_t _pad0;
};
FPT_ABI_ASSERT(sizeof(struct FptVoxelBuildInfo) == 48, "FptVoxelBuildInfo");

int fpt_hip_voxel_build(const struct FptRenderConfig *config,
                        uint32_t *page_table, size_t page_table_len,
                        void *records, size_t records_len,
                        uint16_t *distance, size_t distance_len,
                        uint8_t *material, size_t material_len,
                        float *palette, size_t palette_len,
                        struct FptVoxelBuildInfo *info,
                        char *error,
                        size_t error_len);

/*
 * Exact grid DDA over the voxel tables of `config` for n rays
 * (rays[6*i..6*i+5] = origin xyz, direction xyz): the cells visited in order
 * (traceVoxel, Shaders.metal:4492-4612), up to `cap` per ray, into
 * cells[(i*cap + k)*3 + {0,1,2}]; counts[i] = cells visited (may exceed
 * cap); hits[8*i..] = {hit, entry t, exit t, nx, ny, nz, payload bits,
 * steps} with payload the packed colour (4^3/dense) or material id (8^3).
 */
int fpt_hip_voxel_trace_cells(const struct FptRenderConfig *config,
                              const float *rays,
                              size_t ray_count,
                              uint32_t cap,
                              int32_t *cells,
                              uint32_t *counts,
                              float *hits,
                              char *error,
                              size_t error_len);

/*
 * Batched multi-scene rendering: `count` independent SDF scenes (each its own
 * size, spp, camera and program) rendered by one persistent launch that
 * hands out 8x8 pixel tiles scene after scene from one work counter.  Scene
 * i's image occupies float4 pixels [off_i, off_i + W_i*H_i) of the output,
 * off_i = sum of W_j*H_j for j < i; every scene's image is bit-identical to
 * rendering it alone.  With use_graph (default 1) the launch sequence is
 * captured once per output buffer into a HIP graph and replayed.
 * fpt_hip_batch_render enqueues on `stream` without synchronising.
 */
typedef struct FptHipBatch FptHipBatch;

int fpt_hip_batch_create(const struct FptRenderConfig *configs,
                         uint32_t count,
                         FptHipBatch **batch,
                         char *error,
                         size_t error_len);
size_t fpt_hip_batch_pixels(const FptHipBatch *batch);
int fpt_hip_batch_tune(FptHipBatch *batch,
                       uint32_t workgroups,
                       uint32_t event_batch,
                       uint32_t use_graph,
                       char *error,
                       size_t error_len);
int fpt_hip_batch_render(FptHipBatch *batch,
                         float *device_accum,
                         void *stream,
                         char *error,
                         size_t error_len);
void fpt_hip_batch_destroy(FptHipBatch *batch);

/* Synchronous convenience: render a batch into host memory
 * (linear_output_len >= 4 * sum W_i*H_i floats). */
int fpt_hip_render_batch(const struct FptRenderConfig *configs,
                         uint32_t count,
                         float *linear_output,
                         size_t linear_output_len,
                         double *elapsed_ms,
                         char *error,
                         size_t error_len);

/*
 * Autotuner for exactly one of `plan` / `batch`: launch shapes (with
 * parked paths: workgroups 1-2 per CU, and event_batch left at 0, the
 * launch's default, for a plan, 32..64 for a batch; without: 1-4 per CU x
 * 8..40; both tile orders; for a plan whose proxy is the whole frame, at
 * most 16 samples per pixel, also one slot per pixel against the plan's
 * sample runs) timed
 * on a proxy of the workload (the plan's first samples, up to 16; the batch
 * capped at one sample per pixel) written into device_accum (contents are
 * overwritten), `rounds`
 * rounds in a rotated probe order, median per shape; the fastest shape is
 * stored in the plan/batch and reported.  Synchronous.
 */
struct FptTuneResult {
    uint32_t workgroups;
    uint32_t event_batch;
    uint32_t candidates;
    uint32_t rounds;
    float best_ms;
    float default_ms;   /* median of the default shape (2 per CU, default event batch, longest-first tiles) */
    uint32_t tile_order; /* 1 longest-first, 0 row-major (fpt_hip_set_tile_order) */
    uint32_t sample_groups; /* plan: sample runs per pixel chosen (fpt_hip_plan_tune_stash); batch: 0 */
};
FPT_ABI_ASSERT(sizeof(struct FptTuneResult) == 32, "FptTuneResult");

int fpt_hip_autotune(FptHipPlan *plan,
                     FptHipBatch *batch,
                     float *device_accum,
                     void *stream,
                     uint32_t rounds,
                     struct FptTuneResult *result,
                     char *error,
                     size_t error_len);

/*
 * Order in which a plan's or batch's persistent launch hands out its 8x8
 * tiles (exactly one of plan / batch): 1 (default) longest-first, from one
 * whole path sample per pixel rendered at creation (distance evaluations +
 * orbit iterations, summed per tile), so cheap tiles form the frame's tail
 * and lanes finish together; 0 row-major, scene after scene, which keeps
 * neighbouring tiles together.  Images do not depend on it.  The autotuner
 * tries both.
 */
int fpt_hip_set_tile_order(FptHipPlan *plan, FptHipBatch *batch, uint32_t mode, char *error, size_t error_len);

/*
 * Device plumbing for callers that do not hold a HIP runtime of their own
 * (all on the runtime this library links): device buffers, host<->device
 * copies (to_device 1 = host to device, 0 = device to host, 2 = device to
 * device by a copy kernel: 16-byte aligned pointers and size), a float fill
 * that is safe inside
 * captured graphs, non-blocking streams, and HIP-event timing of `steps`
 * renders of a plan or batch on `stream` (exactly one of plan / batch).
 */
int fpt_hip_device_alloc(size_t bytes, void **ptr, char *error, size_t error_len);
void fpt_hip_device_free(void *ptr);
int fpt_hip_copy(void *dst, const void *src, size_t bytes, uint32_t to_device,
                 char *error, size_t error_len);
int fpt_hip_fill_f32(float *dst, float value, size_t count, void *stream,
                     char *error, size_t error_len);
int fpt_hip_stream_create(void **stream, char *error, size_t error_len);
void fpt_hip_stream_destroy(void *stream);
int fpt_hip_synchronize(void *stream, char *error, size_t error_len);
int fpt_hip_time_renders(FptHipPlan *plan,
                         FptHipBatch *batch,
                         float *device_accum,
                         void *stream,
                         uint32_t steps,
                         double *elapsed_ms,
                         char *error,
                         size_t error_len);

/*
 * Diagnostic image (the fpt_metal_diagnostic_render kernels,
 * Shaders.metal:7282-7442) into host memory: rgba holds width*height*4 bytes
 * in PNG row order (row 0 = top).  *elapsed_ms (optional) is the GPU time of
 * the diagnostic kernel plus the voxel build for the voxel backend.
 */
int fpt_hip_diagnostic_render(const struct FptRenderConfig *config,
                              const struct FptDiagnosticConfig *diagnostic,
                              uint8_t *rgba,
                              size_t rgba_len,
                              double *elapsed_ms,
                              char *error,
                              size_t error_len);

/*
 * Topology specialisation (the reference's runtime-source path,
 * metal_bridge.mm:592-1300 and 4469-4500; validation Shaders.metal:
 * 5055-5103).  A program's opcode sequence is compiled at run time with
 * hipRTC into a straight-line distance / analytic-surface body inside the
 * persistent path tracer; numeric parameters and CSG flags stay in the config,
 * so one compiled module serves every program of that topology.  Modules are
 * cached in memory and on disk ($FPT_HIPRTC_CACHE, default ~/.cache/fpt_hip).
 * A module is attached only after 2^20 validation points agree with the
 * interpreter: |d_s - d_i| <= 2e-5 (1 + |d_i|), gradient error <= 1e-4.
 * fpt_metal_render does the same when config->sdf_topology_specialization is
 * non-zero and reports through its stitch_* out-parameters.
 */
struct FptSpecializationStats {
    uint32_t cache_status;       /* 1 compiled, 2 loaded from a cache */
    uint32_t instruction_count;
    uint64_t source_bytes;
    uint64_t code_bytes;
    uint64_t sample_count;
    uint64_t distance_failures;
    uint64_t gradient_failures;
    float max_distance_error;
    float max_gradient_error;
    double compile_ms;           /* hipRTC compile, or cache read */
    double link_ms;              /* module load */
    double validate_ms;
};
FPT_ABI_ASSERT(sizeof(struct FptSpecializationStats) == 80, "FptSpecializationStats");

/* enable 1: specialise (compile or fetch), validate and attach; enable 0:
 * detach (the plan renders with the interpreter again).  Learned stepping
 * and counting passes keep the interpreter kernel. */
int fpt_hip_plan_specialize(FptHipPlan *plan,
                            uint32_t enable,
                            struct FptSpecializationStats *stats,
                            char *error,
                            size_t error_len);

/* The specialised code itself at caller points (host memory): distance and
 * analytic gradient of the plan's attached module, out_dgrad[4*i..4*i+3] =
 * {distance, d/dx, d/dy, d/dz} as fpt_hip_program_surface reports them for
 * the interpreter.  Fails when no module is attached. */
int fpt_hip_plan_specialized_surface(FptHipPlan *plan,
                                     const float *points_xyzw,
                                     size_t point_count,
                                     float *out_dgrad,
                                     char *error,
                                     size_t error_len);

/* The generated source for `config`'s topology (NUL-terminated; *needed =
 * bytes including the NUL; out may be NULL to query the size). */
int fpt_hip_topology_source(const struct FptRenderConfig *config,
                            char *out,
                            size_t out_len,
                            size_t *needed,
                            char *error,
                            size_t error_len);

/*
 * Learned far-field step proposer.  No counterpart in the reference (its
 * march takes every step on the exact distance, Shaders.metal:3753-3783):
 * a per-scene MLP distance estimate used for march steps far from the
 * surface, evaluated with BF16 matrix instructions.
 *
 * Network: v = p / box_half; 128 Fourier features F[32a + 8q + 2i + c] =
 * sin (c = 0) / cos (c = 1) of pi * 2^((4q + i) / 2) * v_a for axis a < 3,
 * q < 4, i < 4; F[96..98] = v; F[99..127] = 0; then four SiLU layers
 * 128->64->64->64->64 and a linear 64->1 output.  Weights and the inputs of
 * the four hidden layers are rounded to BF16 (nearest even), sums are FP32,
 * the last hidden layer and the output stay FP32.
 *
 * Parameters: FPT_PROPOSER_PARAM_COUNT floats, row-major (unit, input):
 *   W1[64][128] b1[64]  W2[64][64] b2[64]  W3[64][64] b3[64]
 *   W4[64][64] b4[64]   w5[64] b5
 */
#define FPT_PROPOSER_PARAM_COUNT 20801

/* How a plan's marches use the proposal q (all distances in scene units):
 * inside the centred box |x|,|y|,|z| < box_half, a march in far mode steps
 * step_scale * q while q > switch_distance, then continues on the exact
 * distance (backing up one far step if the exact distance there is
 * negative); an exact step with |d| > resume_distance re-enters far mode.
 * Camera marches start in far mode, bounce and shadow marches do not. */
struct FptProposerSettings {
    float box_half;         /* 1.5 for the Mandelbulb training box */
    float switch_distance;  /* 0.02 */
    float resume_distance;  /* >= switch_distance, e.g. 0.1 */
    float step_scale;       /* (0, 1]; 0.5 */
};
FPT_ABI_ASSERT(sizeof(struct FptProposerSettings) == 16, "FptProposerSettings");

/* Learned-stepping work counters of one counting pass. */
struct FptProposerStats {
    uint64_t proposer_calls;  /* wave-wide (64-lane) network evaluations */
    uint64_t proposals;       /* lanes that used a proposal */
    uint64_t far_steps;       /* steps taken on a proposal */
    uint64_t backtracks;      /* far steps undone after crossing the surface */
};
FPT_ABI_ASSERT(sizeof(struct FptProposerStats) == 32, "FptProposerStats");

/*
 * Training samples of `config`'s scene: count float4 (x, y, z, d), d the
 * exact FP32 scene distance.  Points [0, count/2) are uniform in
 * [-box_half, box_half]^3; points [count/2, count) are uniform in the same
 * box restricted to |d| < band (rejection sampling, at most max_attempts
 * candidates per point; *failures counts points that found none and kept
 * their last candidate).  Candidates come from a splitmix64 sequence keyed by
 * (seed, point index), so the set is reproducible.  out_xyzd is host memory.
 */
int fpt_hip_proposer_samples(const struct FptRenderConfig *config,
                             float box_half,
                             float band,
                             uint32_t count,
                             uint64_t seed,
                             uint32_t max_attempts,
                             float *out_xyzd,
                             uint32_t *failures,
                             char *error,
                             size_t error_len);

/* Proposals at points_xyzw[4*i..4*i+2] (host memory) -> out[i]. */
int fpt_hip_proposer_eval(const float *params,
                          size_t param_count,
                          float box_half,
                          const float *points_xyzw,
                          size_t point_count,
                          float *out,
                          char *error,
                          size_t error_len);

/* Attach (settings non-NULL) or detach (NULL) a proposer to an SDF plan;
 * the plan then renders with learned stepping on the persistent kernel. */
int fpt_hip_plan_set_proposer(FptHipPlan *plan,
                              const float *params,
                              size_t param_count,
                              const struct FptProposerSettings *settings,
                              char *error,
                              size_t error_len);

/* Untimed counting pass of a plan's full frame (same samples as a render):
 * the reference's profile categories (Shaders.metal:7650-7704) into stats
 * and, when learned (optional) is non-NULL, the learned-stepping counters.
 * The image goes to device_accum. */
int fpt_hip_plan_profile(FptHipPlan *plan,
                         float *device_accum,
                         struct FptSdfProfileStats *stats,
                         struct FptProposerStats *learned,
                         char *error,
                         size_t error_len);

/*
 * Lane occupancy of the persistent SDF kernel over a plan's full frame (an
 * untimed counting pass, same samples as a render; no counterpart in the
 * reference).  Counts are summed over waves: a "pass" is one execution of a
 * code region by a wave, "lanes" the lanes active in it, so lanes / (64
 * passes) is that region's lane utilisation.
 *   loop_trips          trips of the main state-machine loop
 *   march_passes/lanes  march steps (one distance evaluation, or one orbit
 *                       iteration in the interleaved single-Mandelbulb kernel)
 *   finish_passes/lanes distance finalisations + march steps of the
 *                       interleaved kernel (0 otherwise)
 *   event_passes/lanes  per-hit blocks (hit, shadow end, scatter)
 *   normal_passes/lanes analytic normals inside the per-hit blocks
 *   idle_lanes          lanes without work, summed over loop trips
 *   proposer_lanes      learned-step proposals used (lanes of MLP calls)
 */
struct FptLaneStats {
    uint64_t loop_trips;
    uint64_t march_passes, march_lanes;
    uint64_t finish_passes, finish_lanes;
    uint64_t event_passes, event_lanes;
    uint64_t normal_passes, normal_lanes;
    uint64_t idle_lanes;
    uint64_t proposer_lanes;
    uint64_t _reserved;
};
FPT_ABI_ASSERT(sizeof(struct FptLaneStats) == 96, "FptLaneStats");

int fpt_hip_plan_lane_stats(FptHipPlan *plan,
                            float *device_accum,
                            struct FptLaneStats *stats,
                            char *error,
                            size_t error_len);

/* Memory work of the voxel trace over a plan's full frame: every traversal
 * (camera, bounce and shadow rays) of every sample.  No counterpart in the
 * reference (its profile categories, Shaders.metal:7650-7704, are SDF-only);
 * it gives the voxel DDA's algorithmic bytes: page_reads x 4 B (one
 * page-table entry per brick entered) + cell_reads x the payload read per
 * cell (1 B material id in sparse-bricks8, the 4 B packed colour of a
 * record otherwise). */
struct FptVoxelTraceStats {
    uint64_t visited_cells;   /* DDA steps, hits and misses */
    uint64_t page_reads;
    uint64_t cell_reads;
    uint64_t pixels;
};
FPT_ABI_ASSERT(sizeof(struct FptVoxelTraceStats) == 32, "FptVoxelTraceStats");

/* Untimed counting pass of a voxel plan's full frame (same samples as a
 * render; the image goes to device_accum). */
int fpt_hip_plan_voxel_stats(FptHipPlan *plan,
                             float *device_accum,
                             struct FptVoxelTraceStats *stats,
                             char *error,
                             size_t error_len);

/*
 * Allocation-free batched render.  No counterpart in the reference: this is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It renders
 * `count` SDF scenes exactly as fpt_hip_batch_create + fpt_hip_batch_render
 * do (same kernels, bit-identical images, same output layout), enqueued on
 * `stream` without synchronising, but it calls no hipMalloc, hipFree or other
 * allocating routine: the batch's device state (scene configs, scene table,
 * work counter) lives in a fixed device array inside the library, under 2 MiB,
 * one copy per GPU, held until the process ends.
 *   - count must be 1..FPT_NOALLOC_MAX_SCENES; device_accum_len (floats) must
 *     cover 4 * sum W_i*H_i.
 *   - HDRI-lit configs must already be hydrated (fpt_hip_hydrate_hdri): this
 *     entry point reads no files.
 *   - Every call on one GPU shares that array and keeps the last call's
 *     contents: calls that overlap on one GPU (other threads, or streams that
 *     run concurrently) race.  Serialise them, e.g. on one stream.
 */
#define FPT_NOALLOC_MAX_SCENES 64u
int fpt_hip_render_batch_noalloc(const struct FptRenderConfig *configs,
                                 uint32_t count,
                                 float *device_accum,
                                 size_t device_accum_len,
                                 void *stream,
                                 char *error,
                                 size_t error_len);

/* Last error string of the calling thread (empty when none). */
const char *fpt_hip_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* FPT_ABI_H */
