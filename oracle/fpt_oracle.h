/*
 * ORACLE — test infrastructure only.
 *
 * CPU restatement of the reference's SDF path tracer and voxel DDA
 * (shaders/Shaders.metal, src/metal_bridge.mm, src/mandelbulber/orbit.rs),
 * templated on the arithmetic type: an FP64 mode (the correctness
 * reference) and an FP32 twin (compiled -ffp-contract=off, no fast-math,
 * for bit-exact integer/byte outputs such as DDA cell sequences and brick
 * tables).  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it, and only as the checker.  The product library
 * (libfpt_hip.so) never links it.
 *
 * Parity pins (what each function is checked against) are listed at the top
 * of fpt_oracle.cpp.
 */
#ifndef FPT_ORACLE_H
#define FPT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/fpt_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Precision selector for every entry point: 64 -> double, 32 -> float. */

/* Running-mean float4 accumulation (Shaders.metal:7721-7737) of rows
 * [row_begin, row_end) into accum[(y*W + x)*4]; `threads` OpenMP threads
 * (0 = all).  Returns 0 on success. */
int orc_render(const struct FptRenderConfig *cfg, int precision, int threads,
               uint32_t row_begin, uint32_t row_end, double *accum);

/* Mean of samples [sample_begin, sample_end) per pixel (the RNG keys of
 * sample s are those of sample s of a full render), rows as orc_render. */
int orc_render_samples(const struct FptRenderConfig *cfg, int precision, int threads,
                       uint32_t row_begin, uint32_t row_end, uint32_t sample_begin,
                       uint32_t sample_end, double *accum);

/* Un-jittered primary ray per pixel: march + normal (Shaders.metal:
 * 3753-3783, 3871-3971).  out: 9 doubles per pixel {t, px, py, pz, nx, ny,
 * nz, hit, steps}. */
/* Voxel leaf refinement (Shaders.metal:4312-4490) of one ray interval. */
int orc_leaf_refine(const struct FptRenderConfig *cfg, int precision, uint32_t mode, const double *ray, double *out);
/* environment (Shaders.metal:5823-5844) for n view directions (xyz). */
int orc_environment(const struct FptRenderConfig *cfg, int precision, const double *dirs, size_t n, double *out);
/* estimateFocusDistance (Shaders.metal:5869-5878) of the config's camera. */
double orc_focus_estimate(const struct FptRenderConfig *cfg, int precision);
int orc_trace_primary(const struct FptRenderConfig *cfg, int precision,
                      int threads, double *out);

/* Scene distance (mapSdf) at n points (xyz triples). */
int orc_map_distance(const struct FptRenderConfig *cfg, int precision,
                     const double *points, size_t n, double *out);

/* Scene surface gradient used by normalAt (analytic for typed programs,
 * central differences otherwise) at n points: out = {d, gx, gy, gz}, and
 * the unit normal normalAt() returns: normals = {nx, ny, nz}. */
int orc_camera_ray(const struct FptRenderConfig *cfg, int precision, const uint32_t *xy, size_t n,
                   double *out);
int orc_march(const struct FptRenderConfig *cfg, int precision, const double *rays, size_t n, double *out);
int orc_normal(const struct FptRenderConfig *cfg, int precision,
               const double *points, size_t n, double *normals);

/* Program surface (distance + analytic gradient, Shaders.metal:3266-3369
 * extended through fractal leaves) at n points: out = {d, gx, gy, gz}. */
int orc_program_surface(const struct FptRenderConfig *cfg, int precision,
                        const double *points, size_t n, double *out);

/* One fractal-leaf orbit (opcode + data as in FptSdfInstruction) at a point:
 * out = {distance, radius, derivative, iterations, escaped, zx, zy, zz,
 * trap}. */
int orc_leaf_orbit(uint32_t opcode, const float data[4], int precision,
                   const double point[3], double out[9]);

/* Single formula-kernel steps for the reference's closed-form tests
 * (orbit.rs:1866-1960).  state = {zx, zy, zz, radius, derivative}. */
void orc_mandelbulb_step(double state[5], double power);
void orc_mandelbox_step(double state[5], double scale, double min_r2,
                        double fixed_r2);

/* Mandelbulber native Kaleidoscopic IFS field (Shaders.metal:827-949) at n
 * points: out[4i..4i+3] = {distance, radius, derivative, +-iterations}
 * (negative = escaped). */
int orc_kifs_sample(const struct FptRenderConfig *cfg, int precision, const double *points, size_t n,
                    double *out);

/* Certified distance interval of the typed program over the box centre +-
 * half (programBounds, Shaders.metal:1134-1273): out = {min, max, lipschitz,
 * certified}. */
int orc_program_bounds(const struct FptRenderConfig *cfg, int precision, const double center[3],
                       const double half[3], double out[4]);

/* classifyAccelerationCell (Shaders.metal:4196-4244): out = {interval_min,
 * interval_max, certified, legacy, lipschitz, interval, class}. */
int orc_classify_cell(const struct FptRenderConfig *cfg, int precision, const double center[3],
                      const double half[3], double out[7]);

/* Present (post-process, chromatic aberration, highlight bloom, vertical
 * flip; Shaders.metal:6778-6825, 7462-7507) of a float4 accumulation buffer
 * (width*height*4 floats, row 0 at the top of the scene) into RGBA8. */
int orc_present(const struct FptRenderConfig *cfg, int precision, const float *accum, uint8_t *rgba);

/* Diagnostic image (sdfDiagnostic / voxelDiagnostic, Shaders.metal:
 * 6827-6985, 7210-7442) into RGBA8 in PNG row order.  The voxel backend is
 * FP32 whatever `precision` says. */
int orc_diagnostic(const struct FptRenderConfig *cfg, uint32_t mode, uint32_t target_bounce,
                   float max_distance, int precision, int threads, uint8_t *rgba);

/* hash13 (Shaders.metal:541-555). */
uint32_t orc_hash13_bits(float x, float y, float z);
float orc_hash13(float x, float y, float z);

/* Work counters of one orc_render call (summed over threads):
 * {distance_evals, orbit_iterations, primary_steps, secondary_steps,
 *  shadow_steps, normal_evals, paths, far_steps, far_orbit_iterations};
 * far = march steps whose |distance| exceeded 0.02. */
void orc_last_counters(uint64_t out[9]);

#ifdef __cplusplus
}
#endif

#endif
