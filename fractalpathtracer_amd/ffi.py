"""ctypes mirror of the C ABI in include/fpt_abi.h.

This is the Python counterpart of the reference's src/ffi.rs:1-521: the same
structs (field for field, so sizes match src/ffi.rs:494-507) and the same
entry points, bound to libfpt_hip.so.  Loading fails loudly when the HIP
library has not been built; there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

MAX_OPS = 64
MAX_PRIMS = 64
MAX_STOPS = 16

# FptSdfId (metal_bridge.h:10-24)
SDF_CORNELL_BOX = 0
SDF_GLASS_BALL = 1
SDF_BALL_FRACTAL = 2
SDF_CAGE_FRACTAL = 3
SDF_IFS_FRACTAL = 4
SDF_MANDELBOX_FRACTAL = 5
SDF_MENGER_SPONGE = 6
SDF_TOWER_FRACTAL = 7
SDF_TREE_FRACTAL = 8
SDF_PROGRAM = 9
SDF_README_CORNELL = 10
SDF_README_GLASS = 11
SDF_MANDELBULBER = 12

# FptSdfOpcode (metal_bridge.h:40-56 plus the fractal opcodes of fpt_abi.h)
OP_ABS = 1
OP_TRANSLATE = 2
OP_SCALE = 3
OP_ROTATE_X = 4
OP_ROTATE_Y = 5
OP_ROTATE_Z = 6
OP_REPEAT = 7
OP_SORT_DESC = 8
OP_BOX_FOLD = 9
OP_SPHERE = 16
OP_BOX = 17
OP_PLANE = 18
OP_MANDELBULB = 19
OP_MANDELBOX = 20
OP_MENGER = 21
OP_UNION = 24
OP_INTERSECTION = 25
OP_SUBTRACT = 26
OP_ORBIT_ADD = 32

RENDERER_SDF = 0
RENDERER_VOXEL = 1
SDF_ACCUMULATION_AUTO, SDF_ACCUMULATION_PER_SAMPLE, SDF_ACCUMULATION_BATCH, SDF_ACCUMULATION_CHUNKED = 0, 1, 2, 3

VOXEL_STORAGE_DENSE = 0
VOXEL_STORAGE_SPARSE_BRICKS = 1
VOXEL_STORAGE_SPARSE_BRICKS8 = 8

VOXEL_COVERAGE_LEGACY = 0
VOXEL_COVERAGE_LIPSCHITZ = 1
VOXEL_COVERAGE_INTERVAL = 2
VOXEL_COVERAGE_SHELL = 3

# FptVoxelLeafRefinement (metal_bridge.h, Shaders.metal:45-48).
VOXEL_LEAF_REFINEMENT_NONE = 0
VOXEL_LEAF_REFINEMENT_SECANT_BISECTION = 1
VOXEL_LEAF_REFINEMENT_RESTRICTED_TRACE = 2
VOXEL_LEAF_REFINEMENT_FIXED_DE = 3

# Scene capacity of fpt_hip_render_batch_noalloc (fpt_abi.h).
NOALLOC_MAX_SCENES = 64

u32 = C.c_uint32
f32 = C.c_float


class FptSdfInstruction(C.Structure):
    _fields_ = [("opcode", u32), ("flags", u32), ("material_index", u32), ("_pad0", u32),
                ("data", f32 * 4)]


class FptPrimitiveInstance(C.Structure):
    _fields_ = [("transform", f32 * 12), ("data", f32 * 4), ("opcode", u32),
                ("distance_scale", f32), ("source_instruction", u32), ("_pad0", u32)]


class FptAffineTransform(C.Structure):
    _fields_ = [("transform", f32 * 12), ("distance_scale", f32), ("_pad0", u32 * 3)]


class FptIndexedPrimitive(C.Structure):
    _fields_ = [("data", f32 * 4), ("opcode", u32), ("transform_index", u32),
                ("source_instruction", u32), ("combine_mode", u32)]


class FptTypedSoAProgram(C.Structure):
    _fields_ = [("sphere_count", u32), ("box_count", u32), ("plane_count", u32), ("_pad0", u32)] + [
        (name, (u32 if name.endswith("source") else f32) * MAX_PRIMS)
        for name in ("sphere_x", "sphere_y", "sphere_z", "sphere_radius", "sphere_source",
                     "box_x", "box_y", "box_z", "box_half_x", "box_half_y", "box_half_z", "box_source",
                     "plane_x", "plane_y", "plane_z", "plane_center_x", "plane_center_y",
                     "plane_center_z", "plane_offset", "plane_source")]


class FptRenderConfig(C.Structure):
    _fields_ = [
        ("width", u32), ("height", u32), ("samples", u32), ("preview", u32), ("sdf_id", u32),
        ("glass_mode", u32), ("sdf_accumulation_mode", u32), ("sdf_profile", u32),
        ("sdf_bounce_cap", u32), ("sdf_russian_roulette", u32), ("sdf_normal_mode", u32),
        ("sdf_chunk_samples", u32), ("sdf_rr_start", f32), ("sdf_rr_min_prob", f32),
        ("mandel_iteration_scale", f32),
        ("camera_position", f32 * 3), ("camera_yaw_pitch", f32 * 2), ("camera_roll", f32),
        ("camera_fov", f32), ("camera_dof", f32), ("focus_distance", f32),
        ("render", f32 * 8), ("world", f32 * 7), ("world_one_color", f32 * 3), ("sun", f32 * 5),
        ("sun_color", f32 * 3), ("background_gradient", f32 * 6), ("post", f32 * 7),
        ("set_values", f32 * 40), ("vset_values", f32 * 133),
        ("sdf_program_count", u32), ("gradient_count", u32), ("material_mode", u32),
        ("hdri_enabled", u32), ("hdri_width", u32), ("hdri_height", u32), ("_pad_hdri", u32 * 2),
        ("fractal_style_mode", u32), ("_pad_style", u32 * 3), ("fractal_style", f32 * 12),
        ("program_material", f32 * 8), ("sdf_program", FptSdfInstruction * MAX_OPS),
        ("gradient_stops", (f32 * 4) * MAX_STOPS), ("hdri_path", C.c_uint8 * 256),
        ("hdri_lut", C.c_uint16 * (32 * 16 * 3)),
        ("renderer_backend", u32), ("voxel_resolution", u32), ("voxel_normal_mode", u32),
        ("voxel_storage", u32), ("voxel_bounds_min", f32 * 3), ("voxel_surface_band", f32),
        ("voxel_bounds_max", f32 * 3), ("voxel_fill_interior", u32), ("voxel_coverage_mode", u32),
        ("voxel_build_mode", u32), ("voxel_brick_rejection", u32), ("voxel_leaf_refinement", u32),
        ("voxel_material_mode", u32), ("voxel_offset_mode", u32), ("bound_grid_resolution", u32),
        ("bound_grid_profile", u32), ("bound_grid_profile_stride", u32),
        ("bound_grid_cage_bounds", u32), ("bound_grid_directional", u32),
        ("sdf_program_source_count", u32), ("sdf_program_optimization", u32),
        ("regional_program_resolution", u32), ("bound_grid_fp16", u32),
        ("sdf_flat_union_count", u32),
        ("sdf_flat_union_instances", FptPrimitiveInstance * MAX_PRIMS),
        ("sdf_shading_program_count", u32), ("sdf_geometry_split", u32),
        ("sdf_shading_program", FptSdfInstruction * MAX_OPS),
        ("sdf_topology_specialization", u32), ("sdf_runtime_source_bytecode", u32),
        ("sdf_function_stitching", u32), ("sdf_stitched_surface", u32),
        ("sdf_stitch_validation", u32), ("sdf_stitch_distance_only", u32),
        ("sdf_stitch_split_graph", u32), ("sdf_stitch_fusion", u32), ("sdf_canonical_count", u32),
        ("sdf_canonical_source_count", u32),
        ("sdf_canonical_primitives", FptPrimitiveInstance * MAX_PRIMS),
        ("sdf_canonical_transform_count", u32), ("_pad_canonical_indexed", u32 * 3),
        ("sdf_canonical_transforms", FptAffineTransform * MAX_PRIMS),
        ("sdf_indexed_primitives", FptIndexedPrimitive * MAX_PRIMS),
        ("sdf_typed_soa", FptTypedSoAProgram),
    ]


class FptBoundGridStats(C.Structure):
    _fields_ = [(n, C.c_uint64 * 3) for n in (
        "macro_cells", "certified_skips", "candidate_intervals", "candidate_misses",
        "candidate_hits", "unknown_intervals", "field_evaluations", "directional_steps",
        "cell_exit_clamps", "unknown_derivative_intervals")] + [(n, C.c_uint64) for n in (
            "profiled_paths", "certified_cells", "unknown_cells", "sampled_bound_failures",
            "sampled_false_skips", "certified_derivative_cells", "unknown_derivative_cells",
            "sampled_derivative_failures", "regional_cells", "regional_fallback_cells",
            "regional_unique_programs", "regional_retained_instructions",
            "regional_sampled_distance_failures", "regional_profiled_paths")] + [
        (n, C.c_uint64 * 3) for n in (
            "regional_distance_evaluations", "regional_atlas_evaluations",
            "regional_full_program_evaluations", "regional_cell_entries",
            "regional_same_cell_reuses", "regional_same_program_reuses",
            "regional_program_id_loads", "regional_header_loads",
            "regional_dynamic_instructions")] + [(n, C.c_uint64) for n in (
                "regional_pruned_cells", "regional_pruned_primitives",
                "regional_repeat_seam_fallback_cells", "regional_unsupported_fallback_cells",
                "regional_no_dominance_cells")]


class FptStitchValidationStats(C.Structure):
    _fields_ = [("sample_count", C.c_uint64), ("distance_failures", C.c_uint64),
                ("gradient_failures", C.c_uint64), ("max_distance_error", f32),
                ("max_gradient_error", f32)]


class FptStitchPipelineStats(C.Structure):
    _fields_ = [(n, u32) for n in (
        "instruction_count", "transform_instruction_count", "primitive_instruction_count",
        "primitive_type_runs", "union_count", "intersection_count", "subtraction_count",
        "thread_execution_width", "max_total_threads_per_threadgroup", "graph_node_count")] + [
        ("static_threadgroup_memory_bytes", C.c_uint64), ("runtime_source_bytes", C.c_uint64),
        ("runtime_library_compile_ms", C.c_double), ("runtime_pipeline_link_ms", C.c_double)]


class FptSceneOptions(C.Structure):
    """Render arguments that shape a loaded scene (include/fpt_abi.h)."""
    _fields_ = [(n, u32) for n in (
        "width", "height", "samples", "preview", "sdf_profile", "glass_mode", "program_optimization",
        "geometry_split", "canonical_ir", "flat_union", "typed_soa", "renderer_backend",
        "topology_specialization", "_pad0")]


class FptSdfProfileStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "primary_steps", "secondary_steps", "shadow_steps", "normal_evals", "bounces", "pixels",
        "distance_evals", "march_orbit_iterations", "refinement_steps", "normal_field_evals",
        "material_evals", "max_ray_steps", "max_pixel_steps")] + [
        ("distance_evals_by_phase", C.c_uint64 * 4), ("orbit_iterations_by_phase", C.c_uint64 * 4),
        ("formula_slot_iterations", C.c_uint64 * 9), ("refinement_distance_evals", C.c_uint64)] + [
        (n, C.c_double) for n in ("primary_ms_estimate", "secondary_ms_estimate",
                                  "shadow_ms_estimate", "normal_ms_estimate",
                                  "bounce_ms_estimate")]


class FptAsyncJitStats(C.Structure):
    _fields_ = [("fallback_render_ms", C.c_double), ("jit_build_ms", C.c_double),
                ("stitched_render_ms", C.c_double), ("cache_status", u32),
                ("fallback_completed_before_jit", u32), ("max_absolute_error", f32),
                ("_pad0", u32)]


# FptDiagnosticMode (metal_bridge.h:370-387).
DIAGNOSTIC_MODES = {
    "depth": 0, "normal": 1, "material": 2, "hit_mask": 3, "path_direct": 4, "path_environment": 5,
    "path_throughput": 6, "path_final": 7, "diffuse_normal": 8, "mandel_color_index": 9,
    "mandel_palette_position": 10, "sdf_primary_steps": 11, "sdf_shadow_steps": 12, "sdf_normal_evals": 13,
    "sdf_bounces": 14, "sdf_bounce_contribution": 15,
}


class FptDiagnosticConfig(C.Structure):
    _fields_ = [("mode", u32), ("_pad0", u32), ("max_distance", f32), ("normal_mix", f32),
                ("dispatch_origin", u32 * 2)]


class FptMandelbulberFieldSample(C.Structure):
    _fields_ = [("distance", f32), ("radius", f32), ("derivative", f32), ("iterations", f32)]


class FptRayHit(C.Structure):
    _fields_ = [("t", f32), ("position", f32 * 3), ("normal", f32 * 3), ("hit", u32),
                ("steps", u32), ("_pad0", u32)]


class FptVoxelBuildInfo(C.Structure):
    _fields_ = [("storage", u32), ("brick_size", u32), ("bricks_per_axis", u32), ("active_bricks", u32),
                ("active_cells", C.c_uint64), ("memory_bytes", C.c_uint64), ("build_ms", C.c_double),
                ("rejected_bricks", u32), ("_pad0", u32)]


class FptLaneStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("loop_trips", "march_passes", "march_lanes", "finish_passes",
                                          "finish_lanes", "event_passes", "event_lanes", "normal_passes",
                                          "normal_lanes", "idle_lanes", "proposer_lanes", "_reserved")]


class FptTuneResult(C.Structure):
    _fields_ = [("workgroups", u32), ("event_batch", u32), ("candidates", u32), ("rounds", u32),
                ("best_ms", f32), ("default_ms", f32), ("tile_order", u32), ("sample_groups", u32)]


class FptProposerSettings(C.Structure):
    _fields_ = [("box_half", f32), ("switch_distance", f32), ("resume_distance", f32), ("step_scale", f32)]


class FptProposerStats(C.Structure):
    _fields_ = [("proposer_calls", C.c_uint64), ("proposals", C.c_uint64), ("far_steps", C.c_uint64),
                ("backtracks", C.c_uint64)]


PROPOSER_PARAM_COUNT = 20801


class FptVoxelTraceStats(C.Structure):
    _fields_ = [("visited_cells", C.c_uint64), ("page_reads", C.c_uint64), ("cell_reads", C.c_uint64),
                ("pixels", C.c_uint64)]


class FptSpecializationStats(C.Structure):
    _fields_ = [("cache_status", u32), ("instruction_count", u32), ("source_bytes", C.c_uint64),
                ("code_bytes", C.c_uint64), ("sample_count", C.c_uint64), ("distance_failures", C.c_uint64),
                ("gradient_failures", C.c_uint64), ("max_distance_error", f32), ("max_gradient_error", f32),
                ("compile_ms", C.c_double), ("link_ms", C.c_double), ("validate_ms", C.c_double)]


# Expected sizes (src/ffi.rs:494-507 plus the new FptRayHit).
EXPECTED_SIZES = {
    FptSdfInstruction: 32, FptPrimitiveInstance: 80, FptAffineTransform: 64,
    FptIndexedPrimitive: 32, FptTypedSoAProgram: 5136, FptStitchPipelineStats: 72,
    FptRenderConfig: 30448, FptSdfProfileStats: 288, FptDiagnosticConfig: 24,
    FptMandelbulberFieldSample: 16, FptAsyncJitStats: 40, FptRayHit: 40, FptVoxelBuildInfo: 48,
    FptTuneResult: 32, FptProposerSettings: 16, FptProposerStats: 32, FptSpecializationStats: 80,
    FptSceneOptions: 56, FptVoxelTraceStats: 32, FptLaneStats: 96,
}

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libfpt_hip.so"

# Symbols declared in include/fpt_abi.h.
EXPORTED_SYMBOLS = (
    "fpt_metal_render", "fpt_metal_diagnostic_render", "fpt_mandelbulber_sample_field",
    "fpt_compare_images", "fpt_metal_device_name", "fpt_test_voxel_dda",
    "fpt_test_async_stitch_context", "fpt_test_typed_soa", "fpt_metal_preview",
    "fpt_hip_trace_primary", "fpt_hip_plan_create", "fpt_hip_plan_render",
    "fpt_hip_plan_destroy", "fpt_hip_last_error", "fpt_hip_program_surface", "fpt_hip_voxel_build",
    "fpt_hip_voxel_trace_cells", "fpt_hip_plan_tune", "fpt_hip_plan_tune_stash", "fpt_hip_batch_create", "fpt_hip_batch_pixels",
    "fpt_hip_batch_tune", "fpt_hip_batch_render", "fpt_hip_batch_destroy", "fpt_hip_render_batch",
    "fpt_hip_render_batch_noalloc",
    "fpt_hip_autotune", "fpt_hip_set_tile_order", "fpt_hip_device_alloc", "fpt_hip_device_free", "fpt_hip_copy", "fpt_hip_fill_f32",
    "fpt_hip_stream_create", "fpt_hip_stream_destroy", "fpt_hip_synchronize", "fpt_hip_time_renders",
    "fpt_hip_proposer_samples", "fpt_hip_proposer_eval", "fpt_hip_plan_set_proposer", "fpt_hip_plan_profile",
    "fpt_hip_plan_voxel_stats", "fpt_hip_plan_lane_stats",
    "fpt_hip_diagnostic_render", "fpt_hip_plan_specialize", "fpt_hip_plan_specialized_surface",
    "fpt_hip_topology_source", "fpt_hip_hydrate_hdri",
    "fpt_scene_default_config", "fpt_scene_load", "fpt_scene_load_fpt_settings", "fpt_scene_apply_program_json",
    "fpt_scene_optimize_program", "fpt_scene_split_program", "fpt_scene_compile_canonical",
    "fpt_scene_lower_flat_union", "fpt_scene_lower_typed_soa",
)

_lib = None


def _declare(lib):
    P = C.POINTER
    cp = C.c_char_p
    lib.fpt_metal_render.argtypes = [cp, cp, cp, cp, cp, C.c_size_t, P(FptRenderConfig),
                                     P(C.c_double), P(C.c_double), P(C.c_uint64), P(u32),
                                     P(C.c_uint64), P(u32), P(FptBoundGridStats), P(u32),
                                     P(FptStitchValidationStats), P(FptStitchPipelineStats),
                                     P(FptSdfProfileStats), P(f32), C.c_size_t, C.c_char_p,
                                     C.c_size_t]
    lib.fpt_metal_render.restype = C.c_int
    lib.fpt_metal_diagnostic_render.argtypes = [cp, cp, cp, C.c_size_t, P(FptRenderConfig),
                                                P(FptDiagnosticConfig), P(C.c_double), cp,
                                                C.c_size_t]
    lib.fpt_metal_diagnostic_render.restype = C.c_int
    lib.fpt_mandelbulber_sample_field.argtypes = [cp, cp, C.c_size_t, P(FptRenderConfig),
                                                  P(f32), C.c_size_t,
                                                  P(FptMandelbulberFieldSample), cp, C.c_size_t]
    lib.fpt_mandelbulber_sample_field.restype = C.c_int
    lib.fpt_compare_images.argtypes = [cp, cp, cp, cp, C.c_size_t]
    lib.fpt_compare_images.restype = C.c_int
    lib.fpt_metal_device_name.argtypes = [cp, C.c_size_t]
    lib.fpt_metal_device_name.restype = C.c_int
    lib.fpt_test_voxel_dda.argtypes = [cp, cp, C.c_size_t]
    lib.fpt_test_voxel_dda.restype = C.c_int
    lib.fpt_test_typed_soa.argtypes = [cp, P(FptRenderConfig), P(FptStitchValidationStats), cp, C.c_size_t]
    lib.fpt_test_typed_soa.restype = C.c_int
    lib.fpt_hip_trace_primary.argtypes = [P(FptRenderConfig), P(FptRayHit), C.c_size_t, cp,
                                          C.c_size_t]
    lib.fpt_hip_trace_primary.restype = C.c_int
    lib.fpt_scene_default_config.argtypes = [P(FptRenderConfig)]
    lib.fpt_scene_load.argtypes = [cp, cp, P(FptSceneOptions), P(FptRenderConfig), cp, C.c_size_t, cp, C.c_size_t]
    lib.fpt_scene_load_fpt_settings.argtypes = [cp, P(FptRenderConfig), cp, C.c_size_t]
    lib.fpt_scene_apply_program_json.argtypes = [P(FptRenderConfig), cp, cp, C.c_size_t]
    lib.fpt_scene_optimize_program.argtypes = [P(FptRenderConfig)]
    lib.fpt_scene_optimize_program.restype = None
    lib.fpt_scene_split_program.argtypes = [P(FptRenderConfig), u32]
    lib.fpt_scene_split_program.restype = None
    for name in ("fpt_scene_compile_canonical", "fpt_scene_lower_flat_union", "fpt_scene_lower_typed_soa"):
        getattr(lib, name).argtypes = [P(FptRenderConfig)]
    lib.fpt_hip_hydrate_hdri.argtypes = [P(FptRenderConfig), cp, C.c_size_t]
    lib.fpt_hip_hydrate_hdri.restype = C.c_int
    lib.fpt_hip_plan_create.argtypes = [P(FptRenderConfig), P(C.c_void_p), cp, C.c_size_t]
    lib.fpt_hip_plan_create.restype = C.c_int
    lib.fpt_hip_plan_render.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, cp, C.c_size_t]
    lib.fpt_hip_plan_render.restype = C.c_int
    lib.fpt_hip_plan_destroy.argtypes = [C.c_void_p]
    lib.fpt_hip_plan_destroy.restype = None
    lib.fpt_hip_voxel_build.argtypes = [P(FptRenderConfig), P(u32), C.c_size_t, C.c_void_p, C.c_size_t,
                                         P(C.c_uint16), C.c_size_t, P(C.c_uint8), C.c_size_t, P(f32),
                                         C.c_size_t, P(FptVoxelBuildInfo), cp, C.c_size_t]
    lib.fpt_hip_voxel_build.restype = C.c_int
    lib.fpt_hip_voxel_trace_cells.argtypes = [P(FptRenderConfig), P(f32), C.c_size_t, u32, P(C.c_int32),
                                              P(u32), P(f32), cp, C.c_size_t]
    lib.fpt_hip_voxel_trace_cells.restype = C.c_int
    lib.fpt_hip_program_surface.argtypes = [P(FptRenderConfig), P(f32), C.c_size_t, P(f32), cp, C.c_size_t]
    lib.fpt_hip_program_surface.restype = C.c_int
    lib.fpt_hip_plan_tune.argtypes = [C.c_void_p, u32, u32, u32, cp, C.c_size_t]
    lib.fpt_hip_plan_tune.restype = C.c_int
    lib.fpt_hip_plan_tune_stash.argtypes = [C.c_void_p, u32, u32, cp, C.c_size_t]
    lib.fpt_hip_plan_tune_stash.restype = C.c_int
    lib.fpt_hip_batch_create.argtypes = [P(FptRenderConfig), u32, P(C.c_void_p), cp, C.c_size_t]
    lib.fpt_hip_batch_create.restype = C.c_int
    lib.fpt_hip_batch_pixels.argtypes = [C.c_void_p]
    lib.fpt_hip_batch_pixels.restype = C.c_size_t
    lib.fpt_hip_batch_tune.argtypes = [C.c_void_p, u32, u32, u32, cp, C.c_size_t]
    lib.fpt_hip_batch_tune.restype = C.c_int
    lib.fpt_hip_batch_render.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, cp, C.c_size_t]
    lib.fpt_hip_batch_render.restype = C.c_int
    lib.fpt_hip_batch_destroy.argtypes = [C.c_void_p]
    lib.fpt_hip_batch_destroy.restype = None
    lib.fpt_hip_render_batch.argtypes = [P(FptRenderConfig), u32, P(f32), C.c_size_t, P(C.c_double), cp,
                                         C.c_size_t]
    lib.fpt_hip_render_batch.restype = C.c_int
    lib.fpt_hip_render_batch_noalloc.argtypes = [P(FptRenderConfig), u32, C.c_void_p, C.c_size_t, C.c_void_p, cp,
                                                 C.c_size_t]
    lib.fpt_hip_render_batch_noalloc.restype = C.c_int
    lib.fpt_hip_set_tile_order.argtypes = [C.c_void_p, C.c_void_p, u32, cp, C.c_size_t]
    lib.fpt_hip_set_tile_order.restype = C.c_int
    lib.fpt_hip_autotune.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, u32, P(FptTuneResult), cp,
                                     C.c_size_t]
    lib.fpt_hip_autotune.restype = C.c_int
    lib.fpt_hip_device_alloc.argtypes = [C.c_size_t, P(C.c_void_p), cp, C.c_size_t]
    lib.fpt_hip_device_alloc.restype = C.c_int
    lib.fpt_hip_device_free.argtypes = [C.c_void_p]
    lib.fpt_hip_device_free.restype = None
    lib.fpt_hip_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, u32, cp, C.c_size_t]
    lib.fpt_hip_copy.restype = C.c_int
    lib.fpt_hip_fill_f32.argtypes = [C.c_void_p, f32, C.c_size_t, C.c_void_p, cp, C.c_size_t]
    lib.fpt_hip_fill_f32.restype = C.c_int
    lib.fpt_hip_stream_create.argtypes = [P(C.c_void_p), cp, C.c_size_t]
    lib.fpt_hip_stream_create.restype = C.c_int
    lib.fpt_hip_stream_destroy.argtypes = [C.c_void_p]
    lib.fpt_hip_stream_destroy.restype = None
    lib.fpt_hip_synchronize.argtypes = [C.c_void_p, cp, C.c_size_t]
    lib.fpt_hip_synchronize.restype = C.c_int
    lib.fpt_hip_time_renders.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, u32, P(C.c_double), cp,
                                         C.c_size_t]
    lib.fpt_hip_time_renders.restype = C.c_int
    lib.fpt_hip_proposer_samples.argtypes = [P(FptRenderConfig), f32, f32, u32, C.c_uint64, u32, P(f32), P(u32),
                                             cp, C.c_size_t]
    lib.fpt_hip_proposer_samples.restype = C.c_int
    lib.fpt_hip_proposer_eval.argtypes = [P(f32), C.c_size_t, f32, P(f32), C.c_size_t, P(f32), cp, C.c_size_t]
    lib.fpt_hip_proposer_eval.restype = C.c_int
    lib.fpt_hip_plan_set_proposer.argtypes = [C.c_void_p, P(f32), C.c_size_t, P(FptProposerSettings), cp,
                                              C.c_size_t]
    lib.fpt_hip_plan_set_proposer.restype = C.c_int
    lib.fpt_hip_plan_profile.argtypes = [C.c_void_p, C.c_void_p, P(FptSdfProfileStats), P(FptProposerStats), cp,
                                         C.c_size_t]
    lib.fpt_hip_plan_profile.restype = C.c_int
    lib.fpt_hip_plan_voxel_stats.argtypes = [C.c_void_p, C.c_void_p, P(FptVoxelTraceStats), cp, C.c_size_t]
    lib.fpt_hip_plan_voxel_stats.restype = C.c_int
    lib.fpt_hip_diagnostic_render.argtypes = [P(FptRenderConfig), P(FptDiagnosticConfig), P(C.c_uint8), C.c_size_t,
                                              P(C.c_double), cp, C.c_size_t]
    lib.fpt_hip_diagnostic_render.restype = C.c_int
    lib.fpt_hip_plan_specialized_surface.argtypes = [C.c_void_p, P(f32), C.c_size_t, P(f32), cp, C.c_size_t]
    lib.fpt_hip_plan_specialized_surface.restype = C.c_int
    lib.fpt_hip_plan_lane_stats.argtypes = [C.c_void_p, C.c_void_p, P(FptLaneStats), cp, C.c_size_t]
    lib.fpt_hip_plan_lane_stats.restype = C.c_int
    lib.fpt_hip_plan_specialize.argtypes = [C.c_void_p, u32, P(FptSpecializationStats), cp, C.c_size_t]
    lib.fpt_hip_plan_specialize.restype = C.c_int
    lib.fpt_hip_topology_source.argtypes = [P(FptRenderConfig), cp, C.c_size_t, P(C.c_size_t), cp, C.c_size_t]
    lib.fpt_hip_topology_source.restype = C.c_int
    lib.fpt_hip_last_error.argtypes = []
    lib.fpt_hip_last_error.restype = C.c_char_p


def load_library(path: os.PathLike | None = None):
    """Load libfpt_hip.so (built in-tree by fractalpathtracer_amd.build)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(str(p))
    _declare(lib)
    if path is None:
        _lib = lib
    return lib


class FptError(RuntimeError):
    pass


def error_buffer(n: int = 4096):
    return C.create_string_buffer(n)


def check(status: int, err) -> None:
    if status != 0:
        raise FptError(err.value.decode(errors="replace"))
