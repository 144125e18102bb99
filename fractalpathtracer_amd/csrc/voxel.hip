// Voxel backend for gfx950: sparse-brick voxelizer, exact grid DDA and the
// voxel path tracer.
//
// Reference semantics: Shaders.metal (SM) 3973-4190 (cell record, offsets,
// slab test, cell ownership, next-float DDA advance), 4196-4244
// (classification), 4492-4612 (traceVoxel), 5388-5424 (cell build),
// 6490-6621 (voxel path), 7961-8011 (accumulation); metal_bridge.mm (BR)
// 1718-1816 (staging compaction order) and 5156-5386 (DDA contract).
//
// Every float operation in this file is IEEE single precision with no FMA
// contraction (pragma below) and correctly rounded division/sqrt, so the
// DDA cell sequences, occupancy and brick tables are bit-identical to the
// FP32 CPU restatement.
//
// Storage layouts:
//  * FPT_VOXEL_STORAGE_SPARSE_BRICKS (reference): 4^3 bricks of 12-byte
//    records {packed RGBA8 colour | bit31 occupied, packed unorm4 properties,
//    float emission}; page table u32 per brick, 0 = empty, else page + 1.
//  * FPT_VOXEL_STORAGE_SPARSE_BRICKS8 (new): 8^3 bricks, structure of arrays:
//    FP16 signed distance per cell and a u8 material id per cell (0 = empty,
//    1..255 = palette entry), plus a 256-entry material palette.
//  * FPT_VOXEL_STORAGE_DENSE: res^3 records of the reference layout.
//
// Build: pass 1 classifies every cell, one workgroup per brick, and writes a
// per-brick occupancy flag; pass 2 turns flags into pages with an exclusive
// scan in linear brick order bx + G*(by + G*bz), which is the staging
// compaction order of BR:1744-1770, so page numbers are deterministic; pass 3
// re-evaluates only the occupied bricks and writes their cells.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "fpt_device.h"
#include "post.h"
#include "bounds.h"
#include "voxel_api.h"
#include "persistent_kernel.h"

namespace fpt {

struct VoxelRecord {  // SM:3973-3977
    uint32_t packed_color;
    uint32_t packed_properties;
    float emission;
};
static_assert(sizeof(VoxelRecord) == 12, "VoxelRecord");

// ---------------------------------------------------------------------------
// DDA primitives (SM:4107-4190).
// ---------------------------------------------------------------------------
struct i3 {
    int x, y, z;
};

__device__ __forceinline__ float comp(f3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
__device__ __forceinline__ void setc(f3 &v, int a, float x) {
    if (a == 0) v.x = x; else if (a == 1) v.y = x; else v.z = x;
}
__device__ __forceinline__ int icomp(i3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

__device__ bool ray_aabb(f3 o, f3 d, f3 bmin, f3 bmax, float &near_t, float &far_t, f3 &entry_normal) {
    near_t = -kInf;
    far_t = kInf;
    entry_normal = mk(0.f, 0.f, 0.f);
    for (int a = 0; a < 3; ++a) {
        float da = comp(d, a), oa = comp(o, a);
        if (da == 0.0f) {
            if (oa < comp(bmin, a) || oa > comp(bmax, a)) return false;
            continue;
        }
        float t0 = __fdiv_rn(comp(bmin, a) - oa, da);
        float t1 = __fdiv_rn(comp(bmax, a) - oa, da);
        float an = fminf(t0, t1), af = fmaxf(t0, t1);
        if (an > near_t) {
            near_t = an;
            entry_normal = mk(0.f, 0.f, 0.f);
            setc(entry_normal, a, da > 0.0f ? -1.0f : 1.0f);
        }
        far_t = fminf(far_t, af);
        if (near_t > far_t) return false;
    }
    float entry_t = fmaxf(near_t, 0.0f);
    return isfinite(entry_t) && isfinite(far_t) && far_t >= entry_t;
}

__device__ bool cell_coordinate(f3 pt, f3 dir, f3 bmin, f3 bmax, f3 cs, int res, i3 &cell) {
    int c[3];
    for (int a = 0; a < 3; ++a) {
        float p = comp(pt, a), lo = comp(bmin, a), hi = comp(bmax, a), d = comp(dir, a);
        if (p < lo || p > hi) return false;
        if ((d < 0.0f && p <= lo) || (d > 0.0f && p >= hi)) return false;
    }
    for (int a = 0; a < 3; ++a) {
        float g = __fdiv_rn(comp(pt, a) - comp(bmin, a), comp(cs, a));
        float f = floorf(g);
        bool leaving = comp(dir, a) < 0.0f && g == f;
        int v = int(f) - (leaving ? 1 : 0);
        v = v < res - 1 ? v : res - 1;
        c[a] = v;
    }
    cell = {c[0], c[1], c[2]};
    return c[0] >= 0 && c[1] >= 0 && c[2] >= 0 && c[0] < res && c[1] < res && c[2] < res;
}

__device__ __forceinline__ float next_positive_float(float v) {  // SM:4163-4167
    if (!isfinite(v) || v < 0.0f) return kInf;
    if (v == 0.0f) return __uint_as_float(1u);
    return __uint_as_float(__float_as_uint(v) + 1u);
}

__device__ __forceinline__ bool advance_dda(i3 &cell, i3 step, f3 &next_t, f3 delta_t, float &entry_t,
                                            f3 &entry_normal) {  // SM:4169-4190
    float b = fminf(next_t.x, fminf(next_t.y, next_t.z));
    if (!isfinite(b)) return false;
    bool cx = next_t.x == b, cy = next_t.y == b, cz = next_t.z == b;
    float ne = b;
    if (!(ne > entry_t)) {
        ne = next_positive_float(entry_t);
        if (!isfinite(ne) || !(ne > entry_t)) return false;
    }
    entry_t = ne;
    entry_normal = mk(0.f, 0.f, 0.f);
    int axis = cx ? 0 : (cy ? 1 : 2);
    setc(entry_normal, axis, icomp(step, axis) > 0 ? -1.0f : 1.0f);
    if (cx) { cell.x += step.x; next_t.x = next_t.x + delta_t.x; }
    if (cy) { cell.y += step.y; next_t.y = next_t.y + delta_t.y; }
    if (cz) { cell.z += step.z; next_t.z = next_t.z + delta_t.z; }
    return true;
}

__device__ __forceinline__ f3 cell_size_of(const FptRenderConfig &cfg) {  // SM:4101-4103
    float r = float(cfg.voxel_resolution > 1u ? cfg.voxel_resolution : 1u);
    return mk(__fdiv_rn(cfg.voxel_bounds_max[0] - cfg.voxel_bounds_min[0], r),
              __fdiv_rn(cfg.voxel_bounds_max[1] - cfg.voxel_bounds_min[1], r),
              __fdiv_rn(cfg.voxel_bounds_max[2] - cfg.voxel_bounds_min[2], r));
}

// ---------------------------------------------------------------------------
// Storage access.
// ---------------------------------------------------------------------------
struct VoxelView {
    const uint32_t *page_table;
    const VoxelRecord *records;    // 4^3 layout or dense
    const uint8_t *material;       // 8^3 layout
    const __half *distance;        // 8^3 layout
    const float *palette;          // 8^3 layout: 256 x 8 floats
    uint32_t storage;
    uint32_t brick_shift;          // 2 or 3
    uint32_t brick_grid;
};

struct VHit {
    bool hit;
    float t;        // cell entry distance
    float exit_t;   // cell exit distance
    float surface;  // entry distance, or the leaf refinement's surface distance
    f3 position;    // origin + dir * surface
    f3 cell_center;
    f3 normal;
    uint32_t payload;  // packed colour (4^3/dense) or material id (8^3)
    uint32_t index;    // record index into the pool
    uint32_t steps;    // cells visited up to the hit (0 on a miss, as SM:4602)
    // Memory work of the traversal, hit or miss: cells visited, page-table
    // entries read (once per brick entered) and cell payloads read.
    uint32_t visited, pages, cells;
};

// traceVoxel (SM:4492-4612) without leaf refinement.  `visit` is called with
// every visited cell in order (test hook; a no-op in the renderer).
template <class Visit>
__device__ VHit trace_voxel(f3 origin, f3 dir, const FptRenderConfig &cfg, const VoxelView &v, Visit visit) {
    VHit r;
    r.hit = false;
    r.t = kInf;
    r.exit_t = kInf;
    r.surface = kInf;
    r.steps = 0;
    r.visited = r.pages = r.cells = 0;
    r.payload = 0;
    r.index = 0;
    r.position = r.cell_center = r.normal = mk(0.f, 0.f, 0.f);
    const int res = int(cfg.voxel_resolution > 1u ? cfg.voxel_resolution : 1u);
    const f3 bmin = ld3(cfg.voxel_bounds_min), bmax = ld3(cfg.voxel_bounds_max);
    const f3 cs = cell_size_of(cfg);
    float near_t, far_t;
    f3 entry_normal;
    if (!ray_aabb(origin, dir, bmin, bmax, near_t, far_t, entry_normal)) return r;
    float entry_t = fmaxf(near_t, 0.0f);
    f3 start = origin + dir * entry_t;
    start = mk(clampf(start.x, bmin.x, bmax.x), clampf(start.y, bmin.y, bmax.y), clampf(start.z, bmin.z, bmax.z));
    i3 cell;
    if (!cell_coordinate(start, dir, bmin, bmax, cs, res, cell)) return r;
    i3 step = {0, 0, 0};
    f3 next_t = mk(kInf, kInf, kInf), delta_t = mk(kInf, kInf, kInf);
    for (int a = 0; a < 3; ++a) {
        float da = comp(dir, a);
        int s = 0;
        float nt = kInf, dt = kInf;
        if (da > 0.0f) {
            s = 1;
            float boundary = comp(bmin, a) + float(icomp(cell, a) + 1) * comp(cs, a);
            nt = __fdiv_rn(boundary - comp(origin, a), da);
            dt = __fdiv_rn(comp(cs, a), da);
        } else if (da < 0.0f) {
            s = -1;
            float boundary = comp(bmin, a) + float(icomp(cell, a)) * comp(cs, a);
            nt = __fdiv_rn(boundary - comp(origin, a), da);
            dt = __fdiv_rn(-comp(cs, a), da);
        }
        if (!isfinite(nt) || !isfinite(dt)) {
            nt = kInf;
            dt = kInf;
        }
        if (a == 0) { step.x = s; next_t.x = nt; delta_t.x = dt; }
        else if (a == 1) { step.y = s; next_t.y = nt; delta_t.y = dt; }
        else { step.z = s; next_t.z = nt; delta_t.z = dt; }
    }
    const uint32_t max_steps = uint32_t(res) * 3u + 3u;
    // Page of the current brick, looked up once per brick entered.
    uint32_t cur_brick = 0xffffffffu, page = 0u;
    for (uint32_t it = 0; it < max_steps && entry_t <= far_t; ++it) {
        if (cell.x < 0 || cell.y < 0 || cell.z < 0 || cell.x >= res || cell.y >= res || cell.z >= res) break;
        visit(cell);
        r.visited++;
        uint32_t payload = 0u, index = 0u;
        bool occupied;
        if (v.storage == FPT_VOXEL_STORAGE_DENSE) {
            index = uint32_t(cell.x) + uint32_t(cell.y) * uint32_t(res) + uint32_t(cell.z) * uint32_t(res) * uint32_t(res);
            payload = v.records[index].packed_color;
            r.cells++;
            occupied = (payload & 0x80000000u) != 0u;
        } else {
            const uint32_t sh = v.brick_shift, mask = (1u << sh) - 1u;
            uint32_t bx = uint32_t(cell.x) >> sh, by = uint32_t(cell.y) >> sh, bz = uint32_t(cell.z) >> sh;
            uint32_t brick = bx + by * v.brick_grid + bz * v.brick_grid * v.brick_grid;
            if (brick != cur_brick) {
                cur_brick = brick;
                page = v.page_table[brick];
                r.pages++;
            }
            occupied = false;
            if (page != 0u) {
                r.cells++;
                uint32_t local = (uint32_t(cell.x) & mask) + ((uint32_t(cell.y) & mask) << sh) +
                                 ((uint32_t(cell.z) & mask) << (2u * sh));
                index = ((page - 1u) << (3u * sh)) + local;
                if (v.storage == FPT_VOXEL_STORAGE_SPARSE_BRICKS8) {
                    payload = v.material[index];
                    occupied = payload != 0u;
                } else {
                    payload = v.records[index].packed_color;
                    occupied = (payload & 0x80000000u) != 0u;
                }
            }
        }
        if (occupied) {
            r.hit = true;
            r.t = entry_t;
            r.exit_t = fminf(next_t.x, fminf(next_t.y, next_t.z));
            r.surface = entry_t;
            r.position = origin + dir * entry_t;
            r.cell_center = mk(bmin.x + (float(cell.x) + 0.5f) * cs.x, bmin.y + (float(cell.y) + 0.5f) * cs.y,
                               bmin.z + (float(cell.z) + 0.5f) * cs.z);
            r.normal = entry_normal;
            r.payload = payload;
            r.index = index;
            r.steps = it + 1u;
            return r;
        }
        if (!advance_dda(cell, step, next_t, delta_t, entry_t, entry_normal)) break;
    }
    return r;
}

struct NoVisit {
    __device__ void operator()(i3) const {}
};

// ---------------------------------------------------------------------------
// Leaf refinement of an occupied cell (SM:4312-4490): a surface point inside
// the cell's ray interval [t0, t1], found with the scene's distance.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool sign_differs(float a, float b) { return signbit(a) != signbit(b); }

// Secant/bisection on a signed crossing: a 4-interval scan brackets the
// first sign change, then 8 iterations alternate a guarded secant step (kept
// 10% inside the bracket) with bisection; |d| <= 1e-6 ends early.  Without a
// crossing the hit is left unrefined.
__device__ bool refine_secant_bisection(f3 o, f3 d, float t0, float t1, const FptRenderConfig &cfg, Work &w,
                                        float &out) {
    if (!isfinite(t0) || !isfinite(t1) || !(t1 > t0)) return false;
    const float tol = 1.0e-6f;
    float pt = t0, pd = map_scene(o + d * pt, cfg, w);
    if (!isfinite(pd)) return false;
    if (fabsf(pd) <= tol) {
        out = pt;
        return true;
    }
    float lo_t = t0, hi_t = t1, lo_d = pd, hi_d = pd;
    bool bracket = false;
    for (uint32_t k = 1u; k <= 4u; ++k) {
        const float ct = mixf(t0, t1, float(k) / 4.0f);
        const float cd = map_scene(o + d * ct, cfg, w);
        if (!isfinite(cd)) return false;
        if (fabsf(cd) <= tol) {
            out = ct;
            return true;
        }
        if (sign_differs(pd, cd)) {
            lo_t = pt; lo_d = pd; hi_t = ct; hi_d = cd;
            bracket = true;
            break;
        }
        pt = ct;
        pd = cd;
    }
    if (!bracket) return false;
    for (uint32_t k = 0u; k < 8u; ++k) {
        float ct = 0.5f * (lo_t + hi_t);
        if ((k & 1u) == 0u) {
            const float den = hi_d - lo_d;
            if (isfinite(den) && fabsf(den) > 1.0e-12f) {
                const float st = (lo_t * hi_d - hi_t * lo_d) / den;
                const float guard = (hi_t - lo_t) * 0.1f;
                if (st > lo_t + guard && st < hi_t - guard) ct = st;
            }
        }
        const float cd = map_scene(o + d * ct, cfg, w);
        if (!isfinite(cd)) return false;
        if (fabsf(cd) <= tol) {
            out = ct;
            return true;
        }
        if (sign_differs(lo_d, cd)) {
            hi_t = ct; hi_d = cd;
        } else {
            lo_t = ct; lo_d = cd;
        }
    }
    out = clampf(fabsf(lo_d) <= fabsf(hi_d) ? lo_t : hi_t, t0, t1);
    return true;
}

// Sphere tracing inside the cell for typed programs with certified bounds:
// steps of 0.99 |d| / max(L, 1) (L the cell's Lipschitz bound), at least half
// the tolerance max(1e-6, 1e-4 (t1 - t0)), at most 12, never past the exit
// (where |d| within the tolerance still counts as the surface).
__device__ bool refine_restricted_trace(f3 o, f3 d, float t0, float t1, f3 center, f3 half,
                                        const FptRenderConfig &cfg, Work &w, float &out) {
    if (cfg.sdf_id != FPT_SDF_PROGRAM || !isfinite(t0) || !isfinite(t1) || !(t1 > t0)) return false;
    const ProgramBounds b = program_bounds(center, half, cfg);
    if (!b.certified || !isfinite(b.lipschitz)) return false;
    float t = t0;
    const float tol = fmaxf(1.0e-6f, (t1 - t0) * 1.0e-4f);
    const float inv_l = 1.0f / fmaxf(b.lipschitz, 1.0f);
    for (uint32_t k = 0u; k < 12u; ++k) {
        const float dist = map_scene(o + d * t, cfg, w);
        if (!isfinite(dist)) return false;
        if (fabsf(dist) <= tol) {
            out = clampf(t, t0, t1);
            return true;
        }
        const float nt = t + fmaxf(fabsf(dist) * inv_l * 0.99f, tol * 0.5f);
        if (!(nt < t1)) {
            const float de = map_scene(o + d * t1, cfg, w);
            if (isfinite(de) && fabsf(de) <= tol) {
                out = t1;
                return true;
            }
            return false;
        }
        t = nt;
    }
    return false;
}

// Built-in fractals (unsigned estimators): the smallest |d| of 13 evenly
// spaced samples, then 3 ternary reductions around it; accepted only within
// max(eps, 1e-5, (t1 - t0) / 64) of the surface.
__device__ bool refine_fixed_de(f3 o, f3 d, float t0, float t1, const FptRenderConfig &cfg, Work &w, float &out) {
    const bool fractal = cfg.sdf_id >= FPT_SDF_BALL_FRACTAL && cfg.sdf_id <= FPT_SDF_TREE_FRACTAL;
    if (!fractal || !isfinite(t0) || !isfinite(t1) || !(t1 > t0)) return false;
    const float span = t1 - t0, step = span / 12.0f;
    float best_t = t0, best = kInf;
    for (uint32_t k = 0u; k <= 12u; ++k) {
        const float t = t0 + float(k) * step;
        const float dist = fabsf(map_scene(o + d * t, cfg, w));
        if (!isfinite(dist)) return false;
        if (dist < best) {
            best = dist;
            best_t = t;
        }
    }
    float lo = fmaxf(t0, best_t - step), hi = fminf(t1, best_t + step);
    for (uint32_t k = 0u; k < 3u; ++k) {
        const float lt = mixf(lo, hi, 1.0f / 3.0f), rt = mixf(lo, hi, 2.0f / 3.0f);
        const float ld = fabsf(map_scene(o + d * lt, cfg, w)), rd = fabsf(map_scene(o + d * rt, cfg, w));
        if (!isfinite(ld) || !isfinite(rd)) return false;
        if (ld < best) {
            best = ld;
            best_t = lt;
        }
        if (rd < best) {
            best = rd;
            best_t = rt;
        }
        if (ld <= rd) hi = rt;
        else lo = lt;
    }
    if (best > fmaxf(fmaxf(cfg.render[3], 1.0e-5f), span / 64.0f)) return false;
    out = clampf(best_t, t0, t1);
    return true;
}

// The refinement step of traceVoxel (SM:4583-4602) for a hit h of the ray
// (o, d): moves h.surface and h.position to the refined surface point when
// the configured method finds one.
__device__ void refine_leaf(VHit &h, f3 o, f3 d, const FptRenderConfig &cfg, Work &w) {
    const uint32_t mode = cfg.voxel_leaf_refinement;
    if (!h.hit || mode == FPT_VOXEL_LEAF_REFINEMENT_NONE) return;
    float t = h.t;
    bool ok = false;
    if (mode == FPT_VOXEL_LEAF_REFINEMENT_SECANT_BISECTION)
        ok = refine_secant_bisection(o, d, h.t, h.exit_t, cfg, w, t);
    else if (mode == FPT_VOXEL_LEAF_REFINEMENT_RESTRICTED_TRACE)
        ok = refine_restricted_trace(o, d, h.t, h.exit_t, h.cell_center, cell_size_of(cfg) * 0.5f, cfg, w, t);
    else if (mode == FPT_VOXEL_LEAF_REFINEMENT_FIXED_DE)
        ok = refine_fixed_de(o, d, h.t, h.exit_t, cfg, w, t);
    if (ok) {
        h.surface = t;
        h.position = o + d * t;
    }
}

// ---------------------------------------------------------------------------
// Cell classification and records (SM:4021-4044, 4196-4244, 5388-5414).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t pack_unorm4(float a, float b, float c, float d) {  // SM:4021-4024
    uint32_t q0 = uint32_t(roundf(clampf(a, 0.f, 1.f) * 255.0f));
    uint32_t q1 = uint32_t(roundf(clampf(b, 0.f, 1.f) * 255.0f));
    uint32_t q2 = uint32_t(roundf(clampf(c, 0.f, 1.f) * 255.0f));
    uint32_t q3 = uint32_t(roundf(clampf(d, 0.f, 1.f) * 255.0f));
    return q0 | (q1 << 8u) | (q2 << 16u) | (q3 << 24u);
}

__device__ __forceinline__ f3 unpack_unorm3(uint32_t v) {
    return mk(float(v & 255u) / 255.0f, float((v >> 8u) & 255u) / 255.0f, float((v >> 16u) & 255u) / 255.0f);
}

// Legacy occupancy |d| <= r * max(band, 0.25), or d < 0 with fill_interior.
__device__ __forceinline__ bool classify_occupied(float d, float radius, const FptRenderConfig &cfg) {
    bool occ = isfinite(d) && fabsf(d) <= radius * fmaxf(cfg.voxel_surface_band, 0.25f);
    if (cfg.voxel_fill_interior != 0u && isfinite(d) && d < 0.0f) occ = true;
    return occ;
}

__device__ __forceinline__ f3 cell_center(uint32_t x, uint32_t y, uint32_t z, f3 bmin, f3 cs) {
    return mk(bmin.x + (float(x) + 0.5f) * cs.x, bmin.y + (float(y) + 0.5f) * cs.y, bmin.z + (float(z) + 0.5f) * cs.z);
}

// Material palette id of a parameter: 0 is empty, 1..255 palette entries.
__device__ __forceinline__ uint32_t palette_id(float o, bool constant) {
    if (constant) return 1u;
    float f = o - floorf(o);
    uint32_t bin = uint32_t(floorf(f * 254.0f));
    return 1u + (bin < 253u ? bin : 253u);
}

// ---------------------------------------------------------------------------
// Build kernels.
// ---------------------------------------------------------------------------
// Pass 1: every cell into dense staging (occupancy byte, FP16 distance and
// material id for the 8^3 layout, or the 12-byte record).  One thread per
// cell, x fastest, so staging writes are coalesced.
__global__ __launch_bounds__(256) void voxel_stage_kernel(const FptRenderConfig *__restrict__ cfgp,
                                                          uint8_t *__restrict__ occ, __half *__restrict__ dist,
                                                          uint8_t *__restrict__ mat, VoxelRecord *__restrict__ rec) {
    const FptRenderConfig &cfg = *cfgp;
    const uint32_t res = cfg.voxel_resolution;
    const size_t n = size_t(res) * res * res;
    const size_t i = size_t(blockIdx.x) * 256u + threadIdx.x;
    if (i >= n) return;
    const uint32_t x = uint32_t(i % res), y = uint32_t((i / res) % res), z = uint32_t(i / (size_t(res) * res));
    const f3 cs = cell_size_of(cfg);
    const f3 c = cell_center(x, y, z, ld3(cfg.voxel_bounds_min), cs);
    Work w = {0, 0};
    const float d = distance_scene(c, cfg, w);
    bool o = classify_occupied(d, length(cs * 0.5f), cfg);
    const uint32_t cov = cfg.voxel_coverage_mode;
    if (cov == FPT_VOXEL_COVERAGE_LIPSCHITZ || cov == FPT_VOXEL_COVERAGE_INTERVAL) {  // SM:5388-5398
        const CellClass k = classify_cell(c, cs * 0.5f, d, cfg);
        o = cov == FPT_VOXEL_COVERAGE_LIPSCHITZ ? k.lipschitz : k.interval;
    }
    occ[i] = o ? 1u : 0u;
    if (dist) dist[i] = __float2half_rn(d);
    if (mat) mat[i] = o ? uint8_t(palette_id(material_param(material_orbit(c, cfg, w), cfg), material_is_constant(cfg))) : 0u;
    if (rec) {
        VoxelRecord r = {0u, 0u, 0.0f};
        if (o) {
            Material m = user_material(c, cfg, w);
            r.packed_color = pack_unorm4(m.rgb.x, m.rgb.y, m.rgb.z, 0.0f) | 0x80000000u;
            r.packed_properties = pack_unorm4(m.roughness, m.specular, m.translucency, (m.ior - 1.0f) / 1.5f);
            r.emission = m.emission;
        }
        rec[i] = r;
    }
}

// Occupancy after coverage: shell coverage keeps an occupied cell only if one
// of its 26 neighbours is unoccupied or outside the grid.
__device__ __forceinline__ bool final_occupied(const uint8_t *occ, int x, int y, int z, int res, bool shell) {
    const size_t i = size_t(x) + size_t(y) * res + size_t(z) * res * res;
    if (!occ[i]) return false;
    if (!shell) return true;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                int nx = x + dx, ny = y + dy, nz = z + dz;
                if (nx < 0 || ny < 0 || nz < 0 || nx >= res || ny >= res || nz >= res) return true;
                if (!occ[size_t(nx) + size_t(ny) * res + size_t(nz) * res * res]) return true;
            }
    return false;
}

// Pass 2: one workgroup per brick, flag = any final-occupied cell.
template <int B>
__global__ __launch_bounds__(B * B * B) void voxel_brick_flag_kernel(const uint8_t *__restrict__ occ,
                                                                      uint32_t *__restrict__ flags, uint32_t grid,
                                                                      int res, int shell) {
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    const uint32_t brick = blockIdx.x;
    const uint32_t bx = brick % grid, by = (brick / grid) % grid, bz = brick / (grid * grid);
    const uint32_t l = threadIdx.x;
    const int x = int(bx * B + (l % B)), y = int(by * B + ((l / B) % B)), z = int(bz * B + l / (B * B));
    if (x < res && y < res && z < res && final_occupied(occ, x, y, z, res, shell != 0)) any = 1;
    __syncthreads();
    if (threadIdx.x == 0) flags[brick] = any ? 1u : 0u;
}

// Certified whole-brick rejection of the direct build (SM:5704-5726): one
// thread per brick classifies the brick's box with the program's certified
// bounds; a brick certified outside the surface, or inside it when the
// interior is not filled, keeps no page.  Its cells are then never built, so
// every cell of it is empty whatever the per-cell coverage said.  Brick box:
// first = b*B, end = min(first + B, res), centre bmin + (first + (end -
// first) / 2) * cell, half extent (end - first) * cell / 2, in the
// reference's float order.
template <int B>
__global__ __launch_bounds__(256) void voxel_brick_reject_kernel(const FptRenderConfig *__restrict__ cfgp,
                                                                 uint32_t *__restrict__ flags, uint32_t grid, int res,
                                                                 unsigned long long *__restrict__ rejected) {
    const FptRenderConfig &cfg = *cfgp;
    const uint32_t nb = grid * grid * grid;
    const uint32_t brick = blockIdx.x * 256u + threadIdx.x;
    bool reject = false;
    if (brick < nb) {
        const uint32_t b3[3] = {brick % grid, (brick / grid) % grid, brick / (grid * grid)};
        float c[3], h[3];
        const f3 cs = cell_size_of(cfg);
        const float csa[3] = {cs.x, cs.y, cs.z};
        for (int a = 0; a < 3; ++a) {
            const uint32_t first = b3[a] * uint32_t(B);
            const uint32_t end = first + uint32_t(B) < uint32_t(res) ? first + uint32_t(B) : uint32_t(res);
            const float ext = float(end - first);
            c[a] = cfg.voxel_bounds_min[a] + (float(first) + ext * 0.5f) * csa[a];
            h[a] = ext * csa[a] * 0.5f;
        }
        const f3 center = mk(c[0], c[1], c[2]);
        Work w = {0, 0};
        const CellClass k = classify_cell(center, mk(h[0], h[1], h[2]), distance_scene(center, cfg, w), cfg);
        reject = k.certified && (k.cell_class == CELL_OUTSIDE ||
                                 (cfg.voxel_fill_interior == 0u && k.cell_class == CELL_INSIDE));
        if (reject) flags[brick] = 0u;
    }
    const unsigned long long bal = __ballot(reject);
    if ((threadIdx.x & 63u) == 0u && bal) atomicAdd(rejected, (unsigned long long)__popcll(bal));
}

// Exclusive scan of brick flags -> page table (page + 1, 0 empty), one
// workgroup per 1024 bricks plus a second pass over workgroup totals.
__global__ __launch_bounds__(1024) void scan_local_kernel(const uint32_t *__restrict__ flags, uint32_t *__restrict__ out,
                                                          uint32_t *__restrict__ block_sums, uint32_t n) {
    __shared__ uint32_t s[1024];
    uint32_t i = blockIdx.x * 1024u + threadIdx.x;
    uint32_t v = i < n ? flags[i] : 0u;
    s[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t off = 1; off < 1024u; off <<= 1) {
        uint32_t t = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    if (i < n) out[i] = s[threadIdx.x] - v;  // exclusive within the block
    if (threadIdx.x == 1023u) block_sums[blockIdx.x] = s[1023];
}

__global__ __launch_bounds__(1024) void scan_blocks_kernel(uint32_t *__restrict__ block_sums, uint32_t nblocks,
                                                           uint32_t *__restrict__ total) {
    // Exclusive scan of the block totals by one workgroup, 1024 at a time.
    __shared__ uint32_t s[1024];
    uint32_t carry = 0u;
    for (uint32_t base = 0; base < nblocks; base += 1024u) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < nblocks ? block_sums[i] : 0u;
        s[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t off = 1; off < 1024u; off <<= 1) {
            uint32_t t = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nblocks) block_sums[i] = carry + s[threadIdx.x] - v;
        uint32_t chunk_total = s[1023];
        __syncthreads();
        carry += chunk_total;
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(1024) void page_table_kernel(const uint32_t *__restrict__ flags,
                                                          const uint32_t *__restrict__ local,
                                                          const uint32_t *__restrict__ block_sums,
                                                          uint32_t *__restrict__ page_table, uint32_t n) {
    uint32_t i = blockIdx.x * 1024u + threadIdx.x;
    if (i >= n) return;
    page_table[i] = flags[i] ? block_sums[blockIdx.x] + local[i] + 1u : 0u;
}

// Pass 3: copy the cells of occupied bricks from staging into the pool.
template <int B>
__global__ __launch_bounds__(B * B * B) void voxel_compact_kernel(
    const uint8_t *__restrict__ occ, const __half *__restrict__ sdist, const uint8_t *__restrict__ smat,
    const VoxelRecord *__restrict__ srec, const uint32_t *__restrict__ page_table, uint32_t grid, int res, int shell,
    __half *__restrict__ dist, uint8_t *__restrict__ mat, VoxelRecord *__restrict__ rec,
    unsigned long long *__restrict__ active_cells) {
    const uint32_t brick = blockIdx.x;
    const uint32_t page = page_table[brick];
    if (page == 0u) return;
    const uint32_t bx = brick % grid, by = (brick / grid) % grid, bz = brick / (grid * grid);
    const uint32_t l = threadIdx.x;
    const int x = int(bx * B + (l % B)), y = int(by * B + ((l / B) % B)), z = int(bz * B + l / (B * B));
    const size_t out = size_t(page - 1u) * (B * B * B) + l;
    const bool inside = x < res && y < res && z < res;
    const size_t i = inside ? size_t(x) + size_t(y) * res + size_t(z) * res * res : 0;
    const bool o = inside && final_occupied(occ, x, y, z, res, shell != 0);
    if (B == 8) {
        dist[out] = inside ? sdist[i] : __float2half_rn(kInf);
        mat[out] = o ? smat[i] : 0u;
    } else {
        VoxelRecord r = {0u, 0u, 0.0f};
        if (o) r = srec[i];
        rec[out] = r;
    }
    unsigned long long bal = __ballot(o);
    if ((threadIdx.x & 63u) == 0u && bal) atomicAdd(active_cells, (unsigned long long)__popcll(bal));
}

// Dense storage: drop cells that are not final-occupied (shell coverage) and
// count the rest.
__global__ __launch_bounds__(256) void voxel_dense_finish_kernel(const uint8_t *__restrict__ occ, int res, int shell,
                                                                 VoxelRecord *__restrict__ rec,
                                                                 unsigned long long *__restrict__ active_cells) {
    const size_t n = size_t(res) * res * res;
    const size_t i = size_t(blockIdx.x) * 256u + threadIdx.x;
    bool o = false;
    if (i < n) {
        int x = int(i % res), y = int((i / res) % res), z = int(i / (size_t(res) * res));
        o = final_occupied(occ, x, y, z, res, shell != 0);
        if (!o && occ[i]) rec[i] = VoxelRecord{0u, 0u, 0.0f};
    }
    unsigned long long bal = __ballot(o);
    if ((threadIdx.x & 63u) == 0u && bal) atomicAdd(active_cells, (unsigned long long)__popcll(bal));
}

// Palette of the 8^3 layout: entry id holds the material at the centre of its
// parameter bin (entry 0 unused).
__global__ void voxel_palette_kernel(const FptRenderConfig *__restrict__ cfgp, float *__restrict__ palette) {
    const FptRenderConfig &cfg = *cfgp;
    uint32_t id = threadIdx.x;
    if (id == 0u || id > 255u) return;
    float o = (float(id - 1u) + 0.5f) / 254.0f;
    Material m = material_at_param(o, cfg);
    float *e = palette + id * 8u;
    e[0] = m.rgb.x; e[1] = m.rgb.y; e[2] = m.rgb.z; e[3] = m.roughness;
    e[4] = m.specular; e[5] = m.translucency; e[6] = m.ior; e[7] = m.emission;
}

// ---------------------------------------------------------------------------
// Voxel path tracer (SM:6490-6621).
// ---------------------------------------------------------------------------
// The stored material of a hit cell: palette entry (8^3) or record (4^3,
// dense).
__device__ __forceinline__ Material table_material(const VHit &h, const VoxelView &v) {
    Material m;
    if (v.storage == FPT_VOXEL_STORAGE_SPARSE_BRICKS8) {
        const float *e = v.palette + h.payload * 8u;
        m.rgb = mk(e[0], e[1], e[2]);
        m.roughness = e[3]; m.specular = e[4]; m.translucency = e[5]; m.ior = e[6]; m.emission = e[7];
        return m;
    }
    const VoxelRecord rec = v.records[h.index];  // SM:4033-4044
    m.rgb = unpack_unorm3(rec.packed_color);
    uint32_t p = rec.packed_properties;
    m.roughness = float(p & 255u) / 255.0f;
    m.specular = float((p >> 8u) & 255u) / 255.0f;
    m.translucency = float((p >> 16u) & 255u) / 255.0f;
    m.ior = mixf(1.0f, 2.5f, float((p >> 24u) & 255u) / 255.0f);
    m.emission = rec.emission;
    return m;
}

__device__ Material hit_material(const VHit &h, const FptRenderConfig &cfg, const VoxelView &v, Work &w) {
    if (cfg.voxel_material_mode == FPT_VOXEL_MATERIAL_EXACT) return user_material(h.position, cfg, w);
    return table_material(h, v);
}

// True when a voxel render needs the distance field after the traversal:
// leaf refinement, the exact material or field normals.  Otherwise the
// kernel is built without that code (fewer registers, no call frames).
inline bool voxel_needs_field(const FptRenderConfig &c) {
    return c.voxel_leaf_refinement != FPT_VOXEL_LEAF_REFINEMENT_NONE ||
           c.voxel_material_mode == FPT_VOXEL_MATERIAL_EXACT || c.voxel_normal_mode != 0u;
}

__device__ f3 precision_offset(f3 pos, f3 n) {  // SM:4057-4072 (Ray Tracing Gems)
    const float origin = 1.0f / 32.0f, float_scale = 1.0f / 65536.0f, int_scale = 256.0f;
    f3 r;
    for (int a = 0; a < 3; ++a) {
        float p = comp(pos, a), na = comp(n, a);
        int off = int(int_scale * na);
        int bits = __float_as_int(p);
        bits += p < 0.0f ? -off : off;
        setc(r, a, fabsf(p) < origin ? p + float_scale * na : __int_as_float(bits));
    }
    return r;
}

__device__ f3 offset_origin(f3 pos, f3 n, f3 dir, const FptRenderConfig &cfg) {  // SM:4074-4090
    float side = dot(dir, n) >= 0.0f ? 1.0f : -1.0f;
    f3 dn = n * side;
    if (cfg.voxel_offset_mode == FPT_VOXEL_OFFSET_PRECISION) return precision_offset(pos, dn);
    f3 cs = cell_size_of(cfg);
    float eps = fminf(cs.x, fminf(cs.y, cs.z)) * 0.01f;
    return pos + dn * eps;
}

// Memory work of one thread's traversals (VHit's counters summed).
struct DdaWork {
    uint32_t visited, pages, cells;
    __device__ void add(const VHit &h) {
        visited += h.visited;
        pages += h.pages;
        cells += h.cells;
    }
};

__device__ f3 voxel_sun(f3 pos, UV xy, float seed, const Material &mat, f3 shade_n, f3 off_n,
                        const FptRenderConfig &cfg, const VoxelView &v, int &shadow_steps,
                        DdaWork &dw) {  // SM:6490-6515
    f3 ld = rotate_camera(mk(0.f, 0.f, 1.f), cfg.sun[1] * kPi / 180.0f, cfg.sun[2] * kPi / 180.0f);
    float h1 = hash13(xy.x, xy.y, madd_rn(seed, 5.0f, 1.0f));
    float h2 = hash13(xy.x, xy.y, madd_rn(seed, 3.0f, 5.0f));
    float s, c;
    sincosf(h1 * 2.0f * kPi, &s, &c);
    float m = sqrtf(h2) * cfg.sun[4];
    ld = rotate_camera(ld, c * m, s * m);
    f3 so = offset_origin(pos, off_n, ld, cfg);
    VHit sh = trace_voxel(so, ld, cfg, v, NoVisit());
    shadow_steps += int(sh.steps);
    dw.add(sh);
    if (!sh.hit) {
        float diff = fmaxf(dot(shade_n, ld), 0.0f);
        return ld3(cfg.sun_color) * (cfg.sun[3] * diff * mat.roughness * (1.0f - mat.translucency));
    }
    return mk(0.f, 0.f, 0.f);
}

// FIELD false: the config uses no leaf refinement, a stored material and
// face normals (voxel_needs_field false), so the distance field is never
// evaluated and its code is left out.
template <bool FIELD>
__device__ f3 render_voxel_path(UV xy, uint32_t sample, float focus, const FptRenderConfig &cfg, const VoxelView &v,
                                Work &w, int &steps, DdaWork &dw) {
    const float frame = float(sample);
    f3 pos = ld3(cfg.camera_position);
    const float focal = 1.0f / tanf(cfg.camera_fov / 2.0f * kPi / 180.0f);
    const float aa = 0.3f / fmaxf(float(cfg.width), float(cfg.height));
    float jx, jy;
    random_point(aa, xy, frame, jx, jy);
    f3 dir = rotate_camera_roll(normalize(mk(xy.x + jx, xy.y + jy, focal)), cfg.camera_yaw_pitch[0],
                                cfg.camera_yaw_pitch[1], cfg.camera_roll);
    const f3 fp = pos + dir * focus;
    float lx, ly;
    random_point(cfg.camera_dof, xy, frame, lx, ly);
    pos = pos + rotate_camera_roll(mk(lx, ly, 0.f), cfg.camera_yaw_pitch[0], cfg.camera_yaw_pitch[1], cfg.camera_roll);
    dir = normalize(fp - pos);
    f3 light = mk(0.f, 0.f, 0.f), color = mk(1.f, 1.f, 1.f);
    float sky_mask = 0.0f;
    const float gt = clampf((dir.y + 1.0f) / 2.0f, 0.f, 1.f);
    const f3 gcol = vmix(ld3(&cfg.background_gradient[3]), ld3(&cfg.background_gradient[0]), gt);
    int bounces = int(cfg.render[0]) < 8 ? int(cfg.render[0]) : 8;
    if (cfg.sdf_bounce_cap > 0u && int(cfg.sdf_bounce_cap) < bounces) bounces = int(cfg.sdf_bounce_cap);
    const f3 cs = cell_size_of(cfg);
    const float cmin = fminf(cs.x, fminf(cs.y, cs.z));
    for (int b = 0; b < bounces; ++b) {
        VHit h = trace_voxel(pos, dir, cfg, v, NoVisit());
        if (FIELD) refine_leaf(h, pos, dir, cfg, w);
        steps += int(h.steps);
        dw.add(h);
        if (!h.hit) {
            if (b == 0) sky_mask = 1.0f;
            light = light + environment(dir, cfg);
            break;
        }
        Material mat = FIELD ? hit_material(h, cfg, v, w) : table_material(h, v);
        if (mat.emission > 0.001f) light = light + mat.rgb * mat.emission;
        f3 n = !FIELD ? h.normal
                      : (cfg.voxel_normal_mode == 1u ? normal_scene(h.cell_center, cfg, w)
                                                     : (cfg.voxel_normal_mode == 2u ? normal_scene(h.position, cfg, w)
                                                                                    : h.normal));
        if (dot(n, dir) > 0.0f) n = -n;
        f3 on = cfg.voxel_normal_mode == 2u ? h.normal : n;
        if (dot(on, dir) > 0.0f) on = -on;
        if (cfg.sun[0] == 1.0f) light = light + voxel_sun(h.position, xy, frame, mat, n, on, cfg, v, steps, dw);
        const float fb = float(b);
        float rs = hash13(xy.x, xy.y, madd_rn(frame, 1.37f, fb));
        float rf = hash13(xy.x, xy.y, madd_rn(frame, 7.91f, fb));
        bool transmitted = false;
        if (rs > mat.translucency) {
            f3 refl = reflect(dir, n);
            f3 diff = random_vector(n, xy, madd_rn(frame, 13.37f, fb));
            float q = (mat.ior - 1.0f) / (mat.ior + 1.0f);
            float f0 = q * q;
            float om = 1.0f - clampf(dot(n, -dir), 0.f, 1.f);
            float om2 = om * om;
            float fres = f0 + (1.0f - f0) * (om2 * om2 * om);
            dir = normalize(vmix(refl, diff, mat.roughness));
            if (rf < fres * mat.specular) dir = refl;
            else color = color * mat.rgb;
        } else {
            float q = (mat.ior - 1.0f) / (mat.ior + 1.0f);
            float f0 = q * q;
            float om = 1.0f - clampf(dot(n, -dir), 0.f, 1.f);
            float om2 = om * om;
            float fres = f0 + (1.0f - f0) * (om2 * om2 * om);
            f3 refr = refract(dir, n, 1.0f / mat.ior);
            f3 refl = reflect(dir, n);
            if (dot(refr, refr) < 0.000001f || !isfinite(refr.x) || rf < fres) {
                dir = normalize(refl);
            } else {
                dir = normalize(refr);
                color = color * (vmix(mk(1.f, 1.f, 1.f), mat.rgb, 0.35f) * (1.0f - fres * 0.5f));
                transmitted = true;
            }
        }
        pos = transmitted ? h.position + dir * (cmin * 1.05f) : offset_origin(h.position, on, dir, cfg);
        if (cfg.sdf_russian_roulette != 0u && b + 1 < bounces && float(b + 1) >= cfg.sdf_rr_start) {
            float surv = clampf(fmaxf(color.x, fmaxf(color.y, color.z)), clampf(cfg.sdf_rr_min_prob, 0.01f, 1.0f), 1.0f);
            float rr = hash13(xy.x, xy.y, madd_rn(frame, 19.19f, __fmul_rn(fb, 3.17f)));
            if (rr > surv) break;
            color = color / surv;
        }
    }
    f3 col = light * color;
    if (cfg.world[6] == 1.0f) col = vmix(col, gcol, sky_mask);
    return vmin(col, 8.0f);
}

// estimateVoxelFocusDistance (SM:7630-7636), once per render.
__global__ void voxel_focus_kernel(const FptRenderConfig *__restrict__ cfgp, VoxelView v, float *__restrict__ focus) {
    const FptRenderConfig &cfg = *cfgp;
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (cfg.focus_distance > 0.0f) {
        *focus = cfg.focus_distance;
        return;
    }
    const float focal = 1.0f / tanf(cfg.camera_fov / 2.0f * kPi / 180.0f);
    f3 d = rotate_camera_roll(normalize(mk(0.f, 0.f, focal)), cfg.camera_yaw_pitch[0], cfg.camera_yaw_pitch[1],
                              cfg.camera_roll);
    VHit h = trace_voxel(ld3(cfg.camera_position), d, cfg, v, NoVisit());
    Work w = {0, 0};
    refine_leaf(h, ld3(cfg.camera_position), d, cfg, w);
    *focus = h.hit && h.surface > 0.001f ? h.surface : 5.0f;
}

template <bool FIELD>
__device__ __forceinline__ void voxel_accumulate(const FptRenderConfig *__restrict__ cfgp, const VoxelView &v,
                                                 const float *__restrict__ focus_p, float4 *__restrict__ accum,
                                                 uint32_t start, uint32_t count, unsigned long long *dda) {
    const FptRenderConfig &cfg = *cfgp;
    const uint32_t t = threadIdx.x, wave = t >> 6, lane = t & 63u;
    const uint32_t x = blockIdx.x * 16u + (wave & 1u) * 8u + (lane & 7u);
    const uint32_t y = blockIdx.y * 16u + (wave >> 1) * 8u + (lane >> 3);
    int steps = 0;
    DdaWork dw = {0u, 0u, 0u};
    if (x < cfg.width && y < cfg.height) {
        const float focus = *focus_p;
        const UV uv = screen_uv(x, y, cfg.width, cfg.height);
        const size_t idx = size_t(y) * cfg.width + x;
        f3 c = mk(0.f, 0.f, 0.f);
        if (start != 0u) {
            float4 p = accum[idx];
            c = mk(p.x, p.y, p.z);
        }
        uint32_t total = cfg.samples < 1u ? 1u : (cfg.samples > 512u ? 512u : cfg.samples);
        uint32_t end = start + count < total ? start + count : total;
        Work w = {0, 0};
        for (uint32_t s = start; s < end; ++s) {
            f3 sc = render_voxel_path<FIELD>(uv, s, focus, cfg, v, w, steps, dw);
            c = vmix(c, sc, 1.0f / float(s + 1u));
        }
        accum[idx] = make_float4(c.x, c.y, c.z, 1.0f);
    }
    if (dda) {  // counting pass: {cells visited, page-table reads, payload reads}
        const uint32_t vals[3] = {dw.visited, dw.pages, dw.cells};
        for (int k = 0; k < 3; ++k) {
            unsigned long long s = (unsigned long long)vals[k];
            for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
            if (lane == 0u && s) atomicAdd(&dda[k], s);
        }
    }
}

// Configs that evaluate the distance field after the traversal.
__global__ __launch_bounds__(256) void voxel_accumulate_kernel(const FptRenderConfig *__restrict__ cfgp, VoxelView v,
                                                               const float *__restrict__ focus_p,
                                                               float4 *__restrict__ accum, uint32_t start,
                                                               uint32_t count, unsigned long long *dda) {
    voxel_accumulate<true>(cfgp, v, focus_p, accum, start, count, dda);
}

// Stored materials and face normals only: the traversal's dependent loads
// dominate, so the kernel is held to 8 waves per SIMD (64 VGPRs; the spilled
// registers stay in cache).  Measured on the 1024^3 workload: 4 waves (120
// VGPRs, no spills) 18.7 ms per frame, 8 waves 14.9 ms, images identical.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void voxel_accumulate_table_kernel(
    const FptRenderConfig *__restrict__ cfgp, VoxelView v, const float *__restrict__ focus_p,
    float4 *__restrict__ accum, uint32_t start, uint32_t count, unsigned long long *dda) {
    voxel_accumulate<false>(cfgp, v, focus_p, accum, start, count, dda);
}

// ---------------------------------------------------------------------------
// Persistent voxel path tracer for stored materials and face normals.
//
// The kernel above runs one pixel per lane through all its samples, bounces
// and sun rays, so a wave waits at every traversal for its longest ray (ray
// lengths through the grid range from a few cells to thousands).  Here every
// lane is a state machine over the same path (SM:6526-6621):
//   TRACE  one cell of the exact DDA per loop trip (SM:4560-4611), for
//          camera, bounce and sun rays;
//   EVENT  the traversal has ended; once `event_batch` lanes of the wave
//          wait (or no lane traverses) they shade together: material, sun
//          ray set-up, scatter, path end;
//   IDLE   no pixel left.
// A finished path starts the pixel's next sample in the same lane, a
// finished pixel takes the next 8x8-tile slot from a global counter, so the
// running mean and every RNG key are those of the one-path kernel and the
// image is bit-identical.  BRICKS8: the grid is stored as 8^3 bricks of u8
// material ids (the lookup is specialised); otherwise any layout.
// ---------------------------------------------------------------------------
enum : int { VX_IDLE = 0, VX_TRACE = 1, VX_EVENT = 2 };

// Face normal as a code: 0 = zero vector, 1 + 2*axis + (component > 0).
__device__ __forceinline__ uint32_t normal_code(f3 n) {
    if (n.x != 0.0f) return 1u + (n.x > 0.0f ? 1u : 0u);
    if (n.y != 0.0f) return 3u + (n.y > 0.0f ? 1u : 0u);
    if (n.z != 0.0f) return 5u + (n.z > 0.0f ? 1u : 0u);
    return 0u;
}
__device__ __forceinline__ f3 normal_of(uint32_t c) {
    const float s = (c & 1u) ? -1.0f : 1.0f;  // codes 1, 3, 5 are negative
    return c == 0u ? mk(0.f, 0.f, 0.f) : (c <= 2u ? mk(s, 0.f, 0.f) : (c <= 4u ? mk(0.f, s, 0.f) : mk(0.f, 0.f, s)));
}

// advance_dda with the entry face normal kept as its code (same float
// operations, same cell sequence).
__device__ __forceinline__ bool advance_dda_code(i3 &cell, i3 step, f3 &next_t, f3 delta_t, float &entry_t,
                                                 uint32_t &enc) {  // SM:4169-4190
    float b = fminf(next_t.x, fminf(next_t.y, next_t.z));
    if (!isfinite(b)) return false;
    bool cx = next_t.x == b, cy = next_t.y == b, cz = next_t.z == b;
    float ne = b;
    if (!(ne > entry_t)) {
        ne = next_positive_float(entry_t);
        if (!isfinite(ne) || !(ne > entry_t)) return false;
    }
    entry_t = ne;
    const int axis = cx ? 0 : (cy ? 1 : 2);
    enc = 1u + 2u * uint32_t(axis) + (icomp(step, axis) > 0 ? 0u : 1u);  // normal -1 (code odd) against the step
    if (cx) { cell.x += step.x; next_t.x = next_t.x + delta_t.x; }
    if (cy) { cell.y += step.y; next_t.y = next_t.y + delta_t.y; }
    if (cz) { cell.z += step.z; next_t.z = next_t.z + delta_t.z; }
    return true;
}

#ifndef FPT_VOXEL_PERSISTENT_WAVES
#define FPT_VOXEL_PERSISTENT_WAVES 4
#endif
// DDA cells a traversing lane visits per loop trip.
#ifndef FPT_VOXEL_STEPS_PER_TRIP
#define FPT_VOXEL_STEPS_PER_TRIP 4
#endif

template <bool PROFILE, bool BRICKS8>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FPT_VOXEL_PERSISTENT_WAVES), flatten))
void voxel_persistent_kernel(const FptRenderConfig *__restrict__ cfgp, VoxelView v, const float *__restrict__ focus_p,
                             float4 *__restrict__ accum, uint32_t start, uint32_t count,
                             uint32_t *__restrict__ counter, uint32_t total_slots, int event_batch,
                             unsigned long long *dda) {
    const FptRenderConfig &cfg = *cfgp;
    const uint32_t W = cfg.width, H = cfg.height;
    const uint32_t total = cfg.samples < 1u ? 1u : (cfg.samples > 512u ? 512u : cfg.samples);
    const uint32_t s_end = start + count < total ? start + count : total;
    const float focus = *focus_p;
    const f3 cam = ld3(cfg.camera_position);
    const float focal = 1.0f / tanf(cfg.camera_fov / 2.0f * kPi / 180.0f);
    const float aa = 0.3f / fmaxf(float(W), float(H));
    int bounces = int(cfg.render[0]) < 8 ? int(cfg.render[0]) : 8;
    if (cfg.sdf_bounce_cap > 0u && int(cfg.sdf_bounce_cap) < bounces) bounces = int(cfg.sdf_bounce_cap);
    const int res = int(cfg.voxel_resolution > 1u ? cfg.voxel_resolution : 1u);
    const f3 bmin = ld3(cfg.voxel_bounds_min), bmax = ld3(cfg.voxel_bounds_max);
    const f3 cs = cell_size_of(cfg);
    const float cmin = fminf(cs.x, fminf(cs.y, cs.z));
    const uint32_t max_steps = uint32_t(res) * 3u + 3u;
    const bool sun_on = cfg.sun[0] == 1.0f;
    const f3 sun_dir0 = rotate_camera(mk(0.f, 0.f, 1.f), cfg.sun[1] * kPi / 180.0f, cfg.sun[2] * kPi / 180.0f);

    DdaWork dw = {0u, 0u, 0u};
    // Pixel and path state.
    int phase = VX_IDLE;
    uint32_t pix = 0, s = 0;
    UV uv = {0.f, 0.f};
    f3 acc = mk(0.f, 0.f, 0.f), pos = mk(0.f, 0.f, 0.f), dir = mk(0.f, 0.f, 1.f);
    f3 light = mk(0.f, 0.f, 0.f), color = mk(1.f, 1.f, 1.f), gcol = mk(0.f, 0.f, 0.f);
    float sky_mask = 0.0f;
    int b = 0;
    bool shadow = false;
    // The surface hit a sun ray starts from (kept across the sun ray).
    f3 hpos = mk(0.f, 0.f, 0.f);
    uint32_t hpay = 0u, hidx = 0u, hnc = 0u;
    // Traversal state of the ray in flight (trace_voxel's locals).
    f3 ro = mk(0.f, 0.f, 0.f), rd = mk(0.f, 0.f, 1.f);
    i3 cell = {0, 0, 0}, step = {0, 0, 0};
    f3 next_t = mk(kInf, kInf, kInf), delta_t = mk(kInf, kInf, kInf);
    float entry_t = 0.f, far_t = 0.f;
    uint32_t enc = 0u, it = 0u, cur_brick = 0xffffffffu, page = 0u;
    // Result of the traversal that ended: hit, its entry distance, payload,
    // record index and face normal.
    bool rhit = false;
    float rt = kInf;
    uint32_t rpay = 0u, ridx = 0u, rnc = 0u;

    // traceVoxel's set-up (SM:4492-4559): slab test, entry cell, steps.
    auto start_trace = [&](f3 o, f3 d) __attribute__((always_inline)) {
        ro = o;
        rd = d;
        rhit = false;
        it = 0u;
        cur_brick = 0xffffffffu;
        page = 0u;
        phase = VX_EVENT;  // a ray that misses the grid ends here
        float near_t;
        f3 en;
        if (!ray_aabb(o, d, bmin, bmax, near_t, far_t, en)) return;
        entry_t = fmaxf(near_t, 0.0f);
        f3 st = o + d * entry_t;
        st = mk(clampf(st.x, bmin.x, bmax.x), clampf(st.y, bmin.y, bmax.y), clampf(st.z, bmin.z, bmax.z));
        if (!cell_coordinate(st, d, bmin, bmax, cs, res, cell)) return;
        for (int a = 0; a < 3; ++a) {
            float da = comp(d, a);
            int sg = 0;
            float nt = kInf, dt = kInf;
            if (da > 0.0f) {
                sg = 1;
                float boundary = comp(bmin, a) + float(icomp(cell, a) + 1) * comp(cs, a);
                nt = __fdiv_rn(boundary - comp(o, a), da);
                dt = __fdiv_rn(comp(cs, a), da);
            } else if (da < 0.0f) {
                sg = -1;
                float boundary = comp(bmin, a) + float(icomp(cell, a)) * comp(cs, a);
                nt = __fdiv_rn(boundary - comp(o, a), da);
                dt = __fdiv_rn(-comp(cs, a), da);
            }
            if (!isfinite(nt) || !isfinite(dt)) {
                nt = kInf;
                dt = kInf;
            }
            if (a == 0) { step.x = sg; next_t.x = nt; delta_t.x = dt; }
            else if (a == 1) { step.y = sg; next_t.y = nt; delta_t.y = dt; }
            else { step.z = sg; next_t.z = nt; delta_t.z = dt; }
        }
        enc = normal_code(en);
        phase = VX_TRACE;
    };

    // Camera ray of sample s of the current pixel (render_voxel_path's start).
    auto start_sample = [&]() __attribute__((always_inline)) {
        const float frame = float(s);
        float jx, jy;
        random_point(aa, uv, frame, jx, jy);
        f3 d = rotate_camera_roll(normalize(mk(uv.x + jx, uv.y + jy, focal)), cfg.camera_yaw_pitch[0],
                                  cfg.camera_yaw_pitch[1], cfg.camera_roll);
        const f3 fp = cam + d * focus;
        float lx, ly;
        random_point(cfg.camera_dof, uv, frame, lx, ly);
        pos = cam + rotate_camera_roll(mk(lx, ly, 0.f), cfg.camera_yaw_pitch[0], cfg.camera_yaw_pitch[1],
                                       cfg.camera_roll);
        dir = normalize(fp - pos);
        light = mk(0.f, 0.f, 0.f);
        color = mk(1.f, 1.f, 1.f);
        sky_mask = 0.0f;
        const float gt = clampf((dir.y + 1.0f) / 2.0f, 0.f, 1.f);
        gcol = vmix(ld3(&cfg.background_gradient[3]), ld3(&cfg.background_gradient[0]), gt);
        b = 0;
        shadow = false;
        start_trace(pos, dir);
    };

    bool retry = false;
    auto take_slot = [&](bool want) __attribute__((always_inline)) {
        const uint32_t k = fetch_slot(want, counter);
        retry = false;
        if (!want) return;
        if (k >= total_slots) {
            phase = VX_IDLE;
            return;
        }
        uint32_t x, y;
        pixel_of(k, W, nullptr, x, y);
        if (x >= W || y >= H) {
            retry = true;
            return;
        }
        pix = y * W + x;
        uv = screen_uv(x, y, W, H);
        s = start;
        if (start != 0u) {
            const float4 prev = accum[pix];
            acc = mk(prev.x, prev.y, prev.z);
        } else {
            acc = mk(0.f, 0.f, 0.f);
        }
        start_sample();
    };

    take_slot(true);
    while (__any(retry)) take_slot(retry);

    while (true) {
#pragma unroll 1
        for (int k = 0; k < FPT_VOXEL_STEPS_PER_TRIP && phase == VX_TRACE; ++k) {
            // One iteration of traceVoxel's cell loop (SM:4560-4611).
            bool end = it >= max_steps || !(entry_t <= far_t) || cell.x < 0 || cell.y < 0 || cell.z < 0 ||
                       cell.x >= res || cell.y >= res || cell.z >= res;
            if (!end) {
                dw.visited++;
                uint32_t payload = 0u, index = 0u;
                bool occupied;
                if (BRICKS8) {
                    const uint32_t bx = uint32_t(cell.x) >> 3, by = uint32_t(cell.y) >> 3, bz = uint32_t(cell.z) >> 3;
                    const uint32_t brick = bx + by * v.brick_grid + bz * v.brick_grid * v.brick_grid;
                    if (brick != cur_brick) {
                        cur_brick = brick;
                        page = v.page_table[brick];
                        dw.pages++;
                    }
                    occupied = false;
                    if (page != 0u) {
                        dw.cells++;
                        const uint32_t local = (uint32_t(cell.x) & 7u) + ((uint32_t(cell.y) & 7u) << 3) +
                                               ((uint32_t(cell.z) & 7u) << 6);
                        index = ((page - 1u) << 9) + local;
                        payload = v.material[index];
                        occupied = payload != 0u;
                    }
                } else if (v.storage == FPT_VOXEL_STORAGE_DENSE) {
                    index = uint32_t(cell.x) + uint32_t(cell.y) * uint32_t(res) +
                            uint32_t(cell.z) * uint32_t(res) * uint32_t(res);
                    payload = v.records[index].packed_color;
                    dw.cells++;
                    occupied = (payload & 0x80000000u) != 0u;
                } else {
                    const uint32_t sh = v.brick_shift, mask = (1u << sh) - 1u;
                    uint32_t bx = uint32_t(cell.x) >> sh, by = uint32_t(cell.y) >> sh, bz = uint32_t(cell.z) >> sh;
                    uint32_t brick = bx + by * v.brick_grid + bz * v.brick_grid * v.brick_grid;
                    if (brick != cur_brick) {
                        cur_brick = brick;
                        page = v.page_table[brick];
                        dw.pages++;
                    }
                    occupied = false;
                    if (page != 0u) {
                        dw.cells++;
                        uint32_t local = (uint32_t(cell.x) & mask) + ((uint32_t(cell.y) & mask) << sh) +
                                         ((uint32_t(cell.z) & mask) << (2u * sh));
                        index = ((page - 1u) << (3u * sh)) + local;
                        if (v.storage == FPT_VOXEL_STORAGE_SPARSE_BRICKS8) {
                            payload = v.material[index];
                            occupied = payload != 0u;
                        } else {
                            payload = v.records[index].packed_color;
                            occupied = (payload & 0x80000000u) != 0u;
                        }
                    }
                }
                if (occupied) {
                    rhit = true;
                    rt = entry_t;
                    rpay = payload;
                    ridx = index;
                    rnc = enc;
                    end = true;
                } else {
                    if (!advance_dda_code(cell, step, next_t, delta_t, entry_t, enc)) end = true;
                    ++it;
                }
            }
            if (end) phase = VX_EVENT;
        }
        const unsigned long long ev = __ballot(phase == VX_EVENT);
        const unsigned long long tr = __ballot(phase == VX_TRACE);
        if (!ev && !tr) break;
        if (!(__popcll(ev) >= event_batch || (ev && !tr))) continue;

        bool need_slot = false;
        if (phase == VX_EVENT) {
            const float frame = float(s);
            const float fb = float(b);
            bool path_done = false;
            bool scatter = false;
            f3 n = mk(0.f, 0.f, 0.f);
            Material mat;
            if (!shadow) {
                // End of a camera / bounce traversal (SM:6532-6560).
                if (!rhit) {
                    if (b == 0) sky_mask = 1.0f;
                    light = light + environment(dir, cfg);
                    path_done = true;
                } else {
                    hpos = ro + rd * rt;
                    hpay = rpay;
                    hidx = ridx;
                    hnc = rnc;
                    VHit h;
                    h.payload = hpay;
                    h.index = hidx;
                    mat = table_material(h, v);
                    if (mat.emission > 0.001f) light = light + mat.rgb * mat.emission;
                    if (sun_on) {
                        // Sun ray (SM:6490-6515): its traversal runs as a
                        // TRACE phase; shading resumes at its end.
                        float h1 = hash13(uv.x, uv.y, madd_rn(frame, 5.0f, 1.0f));
                        float h2 = hash13(uv.x, uv.y, madd_rn(frame, 3.0f, 5.0f));
                        float sn, cn;
                        sincosf(h1 * 2.0f * kPi, &sn, &cn);
                        float m = sqrtf(h2) * cfg.sun[4];
                        const f3 ld = rotate_camera(sun_dir0, cn * m, sn * m);
                        n = normal_of(hnc);
                        if (dot(n, dir) > 0.0f) n = -n;
                        shadow = true;
                        start_trace(offset_origin(hpos, n, ld, cfg), ld);
                    } else {
                        scatter = true;
                    }
                }
            } else {
                // End of the sun ray: light it carries (SM:6510-6514).
                shadow = false;
                VHit h;
                h.payload = hpay;
                h.index = hidx;
                mat = table_material(h, v);
                if (!rhit) {
                    n = normal_of(hnc);
                    if (dot(n, dir) > 0.0f) n = -n;
                    float diff = fmaxf(dot(n, rd), 0.0f);
                    light = light + ld3(cfg.sun_color) * (cfg.sun[3] * diff * mat.roughness * (1.0f - mat.translucency));
                }
                scatter = true;
            }
            if (scatter) {
                // Scatter (SM:6561-6613).
                n = normal_of(hnc);
                if (dot(n, dir) > 0.0f) n = -n;
                const f3 on = n;
                float rs = hash13(uv.x, uv.y, madd_rn(frame, 1.37f, fb));
                float rf = hash13(uv.x, uv.y, madd_rn(frame, 7.91f, fb));
                bool transmitted = false;
                if (rs > mat.translucency) {
                    f3 refl = reflect(dir, n);
                    f3 diff = random_vector(n, uv, madd_rn(frame, 13.37f, fb));
                    float q = (mat.ior - 1.0f) / (mat.ior + 1.0f);
                    float f0 = q * q;
                    float om = 1.0f - clampf(dot(n, -dir), 0.f, 1.f);
                    float om2 = om * om;
                    float fres = f0 + (1.0f - f0) * (om2 * om2 * om);
                    dir = normalize(vmix(refl, diff, mat.roughness));
                    if (rf < fres * mat.specular) dir = refl;
                    else color = color * mat.rgb;
                } else {
                    float q = (mat.ior - 1.0f) / (mat.ior + 1.0f);
                    float f0 = q * q;
                    float om = 1.0f - clampf(dot(n, -dir), 0.f, 1.f);
                    float om2 = om * om;
                    float fres = f0 + (1.0f - f0) * (om2 * om2 * om);
                    f3 refr = refract(dir, n, 1.0f / mat.ior);
                    f3 refl = reflect(dir, n);
                    if (dot(refr, refr) < 0.000001f || !isfinite(refr.x) || rf < fres) {
                        dir = normalize(refl);
                    } else {
                        dir = normalize(refr);
                        color = color * (vmix(mk(1.f, 1.f, 1.f), mat.rgb, 0.35f) * (1.0f - fres * 0.5f));
                        transmitted = true;
                    }
                }
                pos = transmitted ? hpos + dir * (cmin * 1.05f) : offset_origin(hpos, on, dir, cfg);
                bool stop = false;
                if (cfg.sdf_russian_roulette != 0u && b + 1 < bounces && float(b + 1) >= cfg.sdf_rr_start) {
                    float surv = clampf(fmaxf(color.x, fmaxf(color.y, color.z)), clampf(cfg.sdf_rr_min_prob, 0.01f, 1.0f),
                                        1.0f);
                    float rr = hash13(uv.x, uv.y, madd_rn(frame, 19.19f, __fmul_rn(fb, 3.17f)));
                    if (rr > surv) stop = true;
                    else color = color / surv;
                }
                ++b;
                if (stop || b >= bounces) path_done = true;
                else start_trace(pos, dir);
            }
            if (path_done) {
                f3 col = light * color;
                if (cfg.world[6] == 1.0f) col = vmix(col, gcol, sky_mask);
                col = vmin(col, 8.0f);
                acc = vmix(acc, col, 1.0f / float(s + 1u));
                ++s;
                if (s < s_end) {
                    start_sample();
                } else {
                    accum[pix] = make_float4(acc.x, acc.y, acc.z, 1.0f);
                    need_slot = true;
                }
            }
        }
        take_slot(need_slot);
        while (__any(retry)) take_slot(retry);
    }

    if (PROFILE && dda) {  // counting pass: {cells visited, page-table reads, payload reads}
        const uint32_t vals[3] = {dw.visited, dw.pages, dw.cells};
        for (int k = 0; k < 3; ++k) {
            unsigned long long sum = (unsigned long long)vals[k];
            for (int off = 32; off > 0; off >>= 1) sum += __shfl_down(sum, off, 64);
            if ((threadIdx.x & 63u) == 0u && sum) atomicAdd(&dda[k], sum);
        }
    }
}

// Test hook: per-ray visit lists and hits for given rays.
struct VisitRecorder {
    int *out;
    uint32_t cap;
    uint32_t *count;
    __device__ void operator()(i3 c) const {
        if (*count < cap) {
            out[*count * 3u + 0u] = c.x;
            out[*count * 3u + 1u] = c.y;
            out[*count * 3u + 2u] = c.z;
        }
        ++*count;
    }
};

__global__ __launch_bounds__(256) void voxel_trace_cells_kernel(const FptRenderConfig *__restrict__ cfgp, VoxelView v,
                                                                const float *__restrict__ rays, size_t n, uint32_t cap,
                                                                int *__restrict__ cells, uint32_t *__restrict__ counts,
                                                                float *__restrict__ hit_out) {
    const FptRenderConfig &cfg = *cfgp;
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *r = rays + i * 6u;
    uint32_t count = 0u;
    VisitRecorder rec{cells + i * size_t(cap) * 3u, cap, &count};
    VHit h = trace_voxel(mk(r[0], r[1], r[2]), mk(r[3], r[4], r[5]), cfg, v, rec);
    counts[i] = count;
    float *o = hit_out + i * 8u;
    o[0] = h.hit ? 1.0f : 0.0f;
    o[1] = h.t;
    o[2] = h.exit_t;
    o[3] = h.normal.x; o[4] = h.normal.y; o[5] = h.normal.z;
    o[6] = __uint_as_float(h.payload);
    o[7] = float(h.steps);
}

// ---------------------------------------------------------------------------
// DDA known-answer contract (SM:5156-5386, the checks that concern the grid
// DDA, slab test, cell ownership and precision offsets).
// ---------------------------------------------------------------------------
// Distances of the 3x3x3 grid over the box lie in its certified bounds
// (programBoundsContainGridSamples, SM:4292-4308).
__device__ bool bounds_contain_grid(f3 c, f3 h, const FptRenderConfig &cfg) {
    const ProgramBounds b = program_bounds(c, h, cfg);
    if (!b.certified || !isfinite(b.minimum) || !isfinite(b.maximum)) return false;
    Work w = {0, 0};
    for (int z = 0; z < 3; ++z)
        for (int y = 0; y < 3; ++y)
            for (int x = 0; x < 3; ++x) {
                const f3 u = mk(float(x) - 1.0f, float(y) - 1.0f, float(z) - 1.0f);
                const float d = distance_scene(c + mk(u.x * h.x, u.y * h.y, u.z * h.z), cfg, w);
                if (!isfinite(d) || d < b.minimum - 2.0e-5f || d > b.maximum + 2.0e-5f) return false;
            }
    return true;
}

__global__ void voxel_dda_contract_kernel(uint32_t *failures_out, const FptRenderConfig *__restrict__ pc) {
    uint32_t f = 0u;
    const f3 bmin = mk(0.f, 0.f, 0.f), bmax = mk(4.f, 4.f, 4.f), cs = mk(1.f, 1.f, 1.f);
    i3 cell = {0, 0, 0};
    auto eq = [](i3 a, int x, int y, int z) { return a.x == x && a.y == y && a.z == z; };
    if (!cell_coordinate(mk(1.f, 2.f, 3.f), mk(1.f, 1.f, 1.f), bmin, bmax, cs, 4, cell) || !eq(cell, 1, 2, 3)) f |= 1u << 0;
    if (!cell_coordinate(mk(1.f, 2.f, 3.f), mk(-1.f, -1.f, -1.f), bmin, bmax, cs, 4, cell) || !eq(cell, 0, 1, 2)) f |= 1u << 1;
    if (!cell_coordinate(mk(4.f, 4.f, 4.f), mk(-1.f, -1.f, -1.f), bmin, bmax, cs, 4, cell) || !eq(cell, 3, 3, 3)) f |= 1u << 2;
    if (cell_coordinate(mk(4.f, 4.f, 4.f), mk(1.f, 1.f, 1.f), bmin, bmax, cs, 4, cell)) f |= 1u << 3;
    if (!cell_coordinate(mk(4.f, 4.f, 4.f), mk(0.f, 0.f, 0.f), bmin, bmax, cs, 4, cell) || !eq(cell, 3, 3, 3)) f |= 1u << 4;
    cell = {0, 0, 0};
    i3 step = {1, 1, 1};
    f3 next_t = mk(1.f, 1.f, 2.f), delta_t = mk(1.f, 1.f, 1.f), en = mk(0.f, 0.f, 0.f);
    float et = 0.0f;
    if (!advance_dda(cell, step, next_t, delta_t, et, en) || !eq(cell, 1, 1, 0) || next_t.x != 2.f || next_t.y != 2.f ||
        next_t.z != 2.f || et != 1.0f || en.x != -1.f || en.y != 0.f || en.z != 0.f)
        f |= 1u << 5;
    if (!advance_dda(cell, step, next_t, delta_t, et, en) || !eq(cell, 2, 2, 1) || next_t.x != 3.f || next_t.y != 3.f ||
        next_t.z != 3.f || et != 2.0f)
        f |= 1u << 6;
    cell = {0, 0, 0};
    next_t = mk(1.f, 2.f, 3.f);
    et = 1.0f;
    if (!advance_dda(cell, step, next_t, delta_t, et, en) || __float_as_uint(et) != __float_as_uint(1.0f) + 1u ||
        !eq(cell, 1, 0, 0))
        f |= 1u << 7;
    float nt = 0.f, ft = 0.f;
    if (!ray_aabb(mk(-1.f, 2.f, 2.f), mk(1.0e-9f, 0.f, 0.f), bmin, bmax, nt, ft, en) || !(nt > 0.0f) || !(ft > nt)) f |= 1u << 8;
    if (ray_aabb(mk(-1.f, 5.f, 2.f), mk(1.f, 0.f, 0.f), bmin, bmax, nt, ft, en)) f |= 1u << 9;
    f3 po = precision_offset(mk(1.f, -1.f, 0.f), mk(1.f, 0.f, 0.f));
    f3 ne = precision_offset(mk(1.f, -1.f, 0.f), mk(-1.f, 0.f, 0.f));
    if (!(po.x > 1.0f) || po.y != -1.0f || !(ne.x < 1.0f) || ne.y != -1.0f) f |= 1u << 27;
    f3 oo = precision_offset(mk(0.f, 0.f, 0.f), mk(1.f, -1.f, 0.f));
    if (!(oo.x > 0.0f) || !(oo.y < 0.0f) || oo.z != 0.0f) f |= 1u << 28;

    // Bounds and cell classification (SM:5211-5298): pc[0..8] are the
    // contract programs (sphere, box, plane, repeat, transformed CSG, ...).
    if (CELL_OUTSIDE != 0u || CELL_INSIDE != 1u || CELL_SURFACE != 2u || CELL_UNKNOWN != 3u) f |= 1u << 10;
    auto cls = [&](f3 c, f3 h, const FptRenderConfig &cfg) {
        Work w = {0, 0};
        return classify_cell(c, h, distance_scene(c, cfg, w), cfg);
    };
    const f3 q = mk(0.25f, 0.25f, 0.25f), z0 = mk(0.f, 0.f, 0.f), t1 = mk(0.1f, 0.1f, 0.1f);
    const ProgramBounds sb = program_bounds(z0, q, pc[0]);
    const CellClass s_in = cls(z0, q, pc[0]), s_out = cls(mk(2.f, 0.f, 0.f), q, pc[0]);
    const CellClass s_surf = cls(mk(1.f, 0.f, 0.f), t1, pc[0]);
    if (!sb.certified || sb.lipschitz != 1.0f || !s_in.certified || s_in.cell_class != CELL_INSIDE ||
        s_out.cell_class != CELL_OUTSIDE || s_surf.cell_class != CELL_SURFACE || !bounds_contain_grid(z0, q, pc[0]))
        f |= 1u << 11;
    if (!s_surf.lipschitz || !s_surf.interval || s_in.lipschitz || s_in.interval) f |= 1u << 20;
    const CellClass miss = cls(mk(0.75f, 0.75f, 0.75f), q, pc[0]);
    if (miss.legacy || !miss.lipschitz || !miss.interval || miss.cell_class != CELL_SURFACE) f |= 1u << 21;
    // Leaf refinement (SM:5309-5341) on the unit sphere: the x axis crosses
    // its surface at t = 1 inside [0.75, 1.25]; the line y = 1.5 misses it.
    {
        Work w = {0, 0};
        const f3 ox = mk(-2.f, 0.f, 0.f), oy = mk(-2.f, 1.5f, 0.f), dx = mk(1.f, 0.f, 0.f), qh = q;
        float rt = 0.0f;
        if (!refine_secant_bisection(ox, dx, 0.75f, 1.25f, pc[0], w, rt) || rt < 0.75f || rt > 1.25f ||
            fabsf(rt - 1.0f) > 1.0e-4f)
            f |= 1u << 22;
        rt = 0.0f;
        if (refine_secant_bisection(oy, dx, 0.75f, 1.25f, pc[0], w, rt)) f |= 1u << 23;
        rt = 0.0f;
        if (!refine_restricted_trace(ox, dx, 0.75f, 1.25f, mk(-1.f, 0.f, 0.f), qh, pc[0], w, rt) || rt < 0.75f ||
            rt > 1.25f || fabsf(rt - 1.0f) > 1.0e-4f)
            f |= 1u << 24;
        rt = 0.0f;
        if (refine_restricted_trace(oy, dx, 0.75f, 1.25f, mk(-1.f, 1.5f, 0.f), qh, pc[0], w, rt)) f |= 1u << 25;
        rt = 0.0f;
        if (refine_fixed_de(ox, dx, 0.75f, 1.25f, pc[0], w, rt)) f |= 1u << 26;  // not a built-in fractal
    }
    if (!bounds_contain_grid(mk(0.35f, -0.15f, 0.2f), mk(0.4f, 0.3f, 0.2f), pc[1])) f |= 1u << 12;
    if (!program_bounds(mk(0.2f, -0.1f, 0.3f), mk(0.4f, 0.2f, 0.5f), pc[2]).certified ||
        !bounds_contain_grid(mk(0.2f, -0.1f, 0.3f), mk(0.4f, 0.2f, 0.5f), pc[2]))
        f |= 1u << 13;
    if (!program_bounds(z0, mk(0.3f, 0.3f, 0.3f), pc[3]).certified ||
        program_bounds(mk(2.f, 0.f, 0.f), t1, pc[3]).certified || !bounds_contain_grid(z0, mk(0.3f, 0.3f, 0.3f), pc[3]))
        f |= 1u << 14;
    if (!bounds_contain_grid(mk(0.4f, -0.2f, 0.1f), mk(0.6f, 0.5f, 0.4f), pc[4])) f |= 1u << 15;
    if (!bounds_contain_grid(mk(0.2f, 0.f, 0.f), mk(0.8f, 0.5f, 0.4f), pc[5])) f |= 1u << 16;
    if (!bounds_contain_grid(mk(0.4f, 0.f, 0.f), mk(1.1f, 0.5f, 0.4f), pc[6])) f |= 1u << 17;
    const CellClass ss = cls(mk(1.f, 0.f, 0.f), t1, pc[7]);
    if (!program_bounds(z0, q, pc[7]).certified || ss.cell_class != CELL_SURFACE || !ss.certified || !ss.legacy ||
        !ss.lipschitz || !ss.interval || !bounds_contain_grid(z0, q, pc[7]))
        f |= 1u << 18;
    if (s_out.legacy || !s_surf.legacy) f |= 1u << 19;
    if (!program_bounds(mk(0.2f, -0.35f, 0.1f), mk(0.7f, 0.45f, 0.55f), pc[8]).certified ||
        !bounds_contain_grid(mk(0.2f, -0.35f, 0.1f), mk(0.7f, 0.45f, 0.55f), pc[8]))
        f |= 1u << 29;
    *failures_out = f;
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
static VoxelView view_of(const VoxelStore &s) {
    VoxelView v;
    v.page_table = static_cast<const uint32_t *>(s.page_table);
    v.records = static_cast<const VoxelRecord *>(s.pool);
    v.material = static_cast<const uint8_t *>(s.material);
    v.distance = static_cast<const __half *>(s.distance);
    v.palette = static_cast<const float *>(s.palette);
    v.storage = s.storage;
    v.brick_shift = s.brick_size == 8 ? 3u : 2u;
    v.brick_grid = s.bricks_per_axis;
    return v;
}

static void seterr(char *error, size_t len, const char *msg, hipError_t e) {
    if (error && len) std::snprintf(error, len, "%s: %s", msg, hipGetErrorString(e));
}

#define VX_TRY(expr, msg)                 \
    do {                                  \
        hipError_t _e = (expr);           \
        if (_e != hipSuccess) {           \
            seterr(error, error_len, msg, _e); \
            for (void *p : temps) (void)hipFree(p); \
            return 1;                     \
        }                                 \
    } while (0)

int voxel_build(const FptRenderConfig &cfg, const FptRenderConfig *dcfg, VoxelStore &store, hipStream_t stream,
                char *error, size_t error_len) {
    std::vector<void *> temps;
    store = VoxelStore();
    const uint32_t res = cfg.voxel_resolution;
    if (res == 0u || res > 2048u) {
        if (error && error_len) std::snprintf(error, error_len, "invalid voxel resolution %u (1..2048)", res);
        return 1;
    }
    const uint32_t storage = cfg.voxel_storage;
    if (storage != FPT_VOXEL_STORAGE_DENSE && storage != FPT_VOXEL_STORAGE_SPARSE_BRICKS &&
        storage != FPT_VOXEL_STORAGE_SPARSE_BRICKS8) {
        if (error && error_len) std::snprintf(error, error_len, "voxel storage %u is not supported", storage);
        return 1;
    }
    if (cfg.voxel_build_mode == FPT_VOXEL_BUILD_DIRECT && storage == FPT_VOXEL_STORAGE_DENSE) {  // BR:4720-4724
        if (error && error_len) std::snprintf(error, error_len, "direct voxel build requires a sparse voxel storage mode");
        return 1;
    }
    const uint32_t cov = cfg.voxel_coverage_mode;
    if (cov > FPT_VOXEL_COVERAGE_SHELL) {
        if (error && error_len) std::snprintf(error, error_len, "voxel coverage mode %u is not supported", cov);
        return 1;
    }
    const int shell = cov == FPT_VOXEL_COVERAGE_SHELL ? 1 : 0;
    store.storage = storage;
    const size_t n = size_t(res) * res * res;
    // The trace addresses cells with 32-bit indices (x + res*(y + res*z) for
    // dense storage, (page-1)*cells_per_brick + local for bricks).
    if (storage == FPT_VOXEL_STORAGE_DENSE && n > size_t(0xffffffffu)) {
        if (error && error_len)
            std::snprintf(error, error_len, "dense voxel resolution %u exceeds 32-bit cell indices (at most 1625)", res);
        return 1;
    }
    unsigned long long *d_cells = nullptr;
    VX_TRY(hipMalloc(&d_cells, sizeof(unsigned long long)), "voxel counter allocation");
    temps.push_back(d_cells);
    VX_TRY(hipMemsetAsync(d_cells, 0, sizeof(unsigned long long), stream), "voxel counter clear");
    uint8_t *occ = nullptr;
    VX_TRY(hipMalloc(&occ, n), "occupancy staging allocation");
    temps.push_back(occ);
    const unsigned nblk = unsigned((n + 255) / 256);
    if (storage == FPT_VOXEL_STORAGE_DENSE) {
        VX_TRY(hipMalloc(&store.pool, n * sizeof(VoxelRecord)), "dense voxel allocation");
        hipLaunchKernelGGL(voxel_stage_kernel, dim3(nblk), dim3(256), 0, stream, dcfg, occ, nullptr, nullptr,
                           static_cast<VoxelRecord *>(store.pool));
        VX_TRY(hipGetLastError(), "voxel stage kernel");
        hipLaunchKernelGGL(voxel_dense_finish_kernel, dim3(nblk), dim3(256), 0, stream, occ, int(res), shell,
                           static_cast<VoxelRecord *>(store.pool), d_cells);
        VX_TRY(hipGetLastError(), "voxel dense kernel");
        VX_TRY(hipMalloc(&store.page_table, 4), "page table allocation");
        VX_TRY(hipMemsetAsync(store.page_table, 0, 4, stream), "page table clear");
        unsigned long long cells = 0;
        VX_TRY(hipMemcpyAsync(&cells, d_cells, sizeof(cells), hipMemcpyDeviceToHost, stream), "counter readback");
        VX_TRY(hipStreamSynchronize(stream), "dense build");
        store.active_cells = cells;
        store.memory_bytes = n * sizeof(VoxelRecord) + 4;
        for (void *p : temps) (void)hipFree(p);
        return 0;
    }
    const uint32_t B = storage == FPT_VOXEL_STORAGE_SPARSE_BRICKS8 ? 8u : 4u;
    const uint32_t G = (res + B - 1u) / B;
    const uint32_t nb = G * G * G;
    // The direct build (metal_bridge.mm:1818-1870) may reject whole bricks
    // whose certified program bounds exclude the surface (SM:5704-5726).
    const bool reject_bricks = cfg.voxel_build_mode == FPT_VOXEL_BUILD_DIRECT && cfg.voxel_brick_rejection != 0u &&
                               cfg.sdf_id == FPT_SDF_PROGRAM && cfg.sdf_program_count > 0u;
    store.brick_size = B;
    store.bricks_per_axis = G;
    __half *sdist = nullptr;
    uint8_t *smat = nullptr;
    VoxelRecord *srec = nullptr;
    if (B == 8u) {
        VX_TRY(hipMalloc(&sdist, n * sizeof(__half)), "distance staging allocation");
        temps.push_back(sdist);
        VX_TRY(hipMalloc(&smat, n), "material staging allocation");
        temps.push_back(smat);
    } else {
        VX_TRY(hipMalloc(&srec, n * sizeof(VoxelRecord)), "record staging allocation");
        temps.push_back(srec);
    }
    hipLaunchKernelGGL(voxel_stage_kernel, dim3(nblk), dim3(256), 0, stream, dcfg, occ, sdist, smat, srec);
    VX_TRY(hipGetLastError(), "voxel stage kernel");
    uint32_t *flags = nullptr, *local = nullptr, *sums = nullptr, *total = nullptr;
    const uint32_t nblocks = (nb + 1023u) / 1024u;
    VX_TRY(hipMalloc(&flags, size_t(nb) * 4), "brick flag allocation");
    temps.push_back(flags);
    VX_TRY(hipMalloc(&local, size_t(nb) * 4), "scan allocation");
    temps.push_back(local);
    VX_TRY(hipMalloc(&sums, size_t(nblocks) * 4), "scan allocation");
    temps.push_back(sums);
    VX_TRY(hipMalloc(&total, 4), "scan allocation");
    temps.push_back(total);
    VX_TRY(hipMalloc(&store.page_table, size_t(nb) * 4), "page table allocation");
    if (B == 8u)
        hipLaunchKernelGGL(voxel_brick_flag_kernel<8>, dim3(nb), dim3(512), 0, stream, occ, flags, G, int(res), shell);
    else
        hipLaunchKernelGGL(voxel_brick_flag_kernel<4>, dim3(nb), dim3(64), 0, stream, occ, flags, G, int(res), shell);
    VX_TRY(hipGetLastError(), "brick flag kernel");
    unsigned long long *d_rejected = nullptr;
    if (reject_bricks) {
        VX_TRY(hipMalloc(&d_rejected, sizeof(unsigned long long)), "rejection counter allocation");
        temps.push_back(d_rejected);
        VX_TRY(hipMemsetAsync(d_rejected, 0, sizeof(unsigned long long), stream), "rejection counter clear");
        if (B == 8u)
            hipLaunchKernelGGL(voxel_brick_reject_kernel<8>, dim3((nb + 255u) / 256u), dim3(256), 0, stream, dcfg,
                               flags, G, int(res), d_rejected);
        else
            hipLaunchKernelGGL(voxel_brick_reject_kernel<4>, dim3((nb + 255u) / 256u), dim3(256), 0, stream, dcfg,
                               flags, G, int(res), d_rejected);
        VX_TRY(hipGetLastError(), "brick rejection kernel");
    }
    hipLaunchKernelGGL(scan_local_kernel, dim3(nblocks), dim3(1024), 0, stream, flags, local, sums, nb);
    VX_TRY(hipGetLastError(), "scan kernel");
    hipLaunchKernelGGL(scan_blocks_kernel, dim3(1), dim3(1024), 0, stream, sums, nblocks, total);
    VX_TRY(hipGetLastError(), "scan kernel");
    hipLaunchKernelGGL(page_table_kernel, dim3(nblocks), dim3(1024), 0, stream, flags, local, sums,
                       static_cast<uint32_t *>(store.page_table), nb);
    VX_TRY(hipGetLastError(), "page table kernel");
    uint32_t pages = 0;
    unsigned long long rejected = 0;
    VX_TRY(hipMemcpyAsync(&pages, total, 4, hipMemcpyDeviceToHost, stream), "page count readback");
    if (reject_bricks)
        VX_TRY(hipMemcpyAsync(&rejected, d_rejected, sizeof(rejected), hipMemcpyDeviceToHost, stream),
               "rejection count readback");
    VX_TRY(hipStreamSynchronize(stream), "voxel classification");
    if (reject_bricks && B == 4u) {
        // The reference's direct build holds at most page_capacity 4^3 bricks
        // (480 MiB at resolutions >= 512, metal_bridge.mm:1824-1834) and
        // falls back to the staging build, which rejects nothing, when the
        // kept bricks do not fit (metal_bridge.mm:5411-5440).
        const uint64_t brick_bytes = 64u * sizeof(VoxelRecord);
        const uint64_t page_bytes = uint64_t(nb) * sizeof(uint32_t);
        const uint64_t budget = res >= 512u ? 480ull * 1024u * 1024u : page_bytes + uint64_t(nb) * brick_bytes;
        const uint64_t available = budget > page_bytes ? budget - page_bytes : brick_bytes;
        const uint64_t capacity = std::max<uint64_t>(1u, std::min<uint64_t>(nb, available / brick_bytes));
        if (pages > capacity) {
            rejected = 0;
            hipLaunchKernelGGL(voxel_brick_flag_kernel<4>, dim3(nb), dim3(64), 0, stream, occ, flags, G, int(res),
                               shell);
            hipLaunchKernelGGL(scan_local_kernel, dim3(nblocks), dim3(1024), 0, stream, flags, local, sums, nb);
            hipLaunchKernelGGL(scan_blocks_kernel, dim3(1), dim3(1024), 0, stream, sums, nblocks, total);
            hipLaunchKernelGGL(page_table_kernel, dim3(nblocks), dim3(1024), 0, stream, flags, local, sums,
                               static_cast<uint32_t *>(store.page_table), nb);
            VX_TRY(hipGetLastError(), "staging fallback kernels");
            VX_TRY(hipMemcpyAsync(&pages, total, 4, hipMemcpyDeviceToHost, stream), "page count readback");
            VX_TRY(hipStreamSynchronize(stream), "staging fallback");
        }
    }
    store.rejected_bricks = uint32_t(rejected);
    store.active_bricks = pages;
    const size_t cpb = size_t(B) * B * B;
    if (size_t(pages) * cpb > size_t(0xffffffffu)) {
        if (error && error_len)
            std::snprintf(error, error_len, "%u occupied bricks exceed 32-bit cell indices of the brick pool", pages);
        for (void *p : temps) (void)hipFree(p);
        voxel_free(store);
        return 1;
    }
    const size_t pool_cells = size_t(pages > 0u ? pages : 1u) * cpb;
    if (B == 8u) {
        VX_TRY(hipMalloc(&store.distance, pool_cells * sizeof(__half)), "distance pool allocation");
        VX_TRY(hipMalloc(&store.material, pool_cells), "material pool allocation");
        VX_TRY(hipMalloc(&store.palette, 256 * 8 * sizeof(float)), "palette allocation");
        VX_TRY(hipMemsetAsync(store.palette, 0, 256 * 8 * sizeof(float), stream), "palette clear");
        hipLaunchKernelGGL(voxel_palette_kernel, dim3(1), dim3(256), 0, stream, dcfg, static_cast<float *>(store.palette));
        VX_TRY(hipGetLastError(), "palette kernel");
        if (pages)
            hipLaunchKernelGGL(voxel_compact_kernel<8>, dim3(nb), dim3(512), 0, stream, occ, sdist, smat, nullptr,
                               static_cast<const uint32_t *>(store.page_table), G, int(res), shell,
                               static_cast<__half *>(store.distance), static_cast<uint8_t *>(store.material), nullptr,
                               d_cells);
        store.memory_bytes = size_t(nb) * 4 + pool_cells * 3 + 256 * 8 * sizeof(float);
    } else {
        VX_TRY(hipMalloc(&store.pool, pool_cells * sizeof(VoxelRecord)), "brick pool allocation");
        if (pages)
            hipLaunchKernelGGL(voxel_compact_kernel<4>, dim3(nb), dim3(64), 0, stream, occ, nullptr, nullptr, srec,
                               static_cast<const uint32_t *>(store.page_table), G, int(res), shell, nullptr, nullptr,
                               static_cast<VoxelRecord *>(store.pool), d_cells);
        store.memory_bytes = size_t(nb) * 4 + pool_cells * sizeof(VoxelRecord);
    }
    VX_TRY(hipGetLastError(), "voxel compact kernel");
    unsigned long long cells = 0;
    VX_TRY(hipMemcpyAsync(&cells, d_cells, sizeof(cells), hipMemcpyDeviceToHost, stream), "counter readback");
    VX_TRY(hipStreamSynchronize(stream), "voxel build");
    store.active_cells = cells;
    for (void *p : temps) (void)hipFree(p);
    return 0;
}

void voxel_free(VoxelStore &s) {
    for (void *p : {s.page_table, s.pool, s.material, s.distance, s.palette, s.focus, s.counter})
        if (p) (void)hipFree(p);
    s = VoxelStore();
}

hipError_t launch_voxel_accumulate(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg, VoxelStore &store,
                                   float4 *accum, uint32_t start, uint32_t count, hipStream_t stream,
                                   unsigned long long *dda, bool persistent, uint32_t workgroups, int event_batch) {
    VoxelView v = view_of(store);
    if (!store.focus) {
        hipError_t e = hipMalloc(&store.focus, sizeof(float));
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(voxel_focus_kernel, dim3(1), dim3(64), 0, stream, dcfg, v, static_cast<float *>(store.focus));
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (persistent && !voxel_needs_field(hcfg)) {
        if (!store.counter) {
            hipError_t e = hipMalloc(&store.counter, sizeof(uint32_t));
            if (e != hipSuccess) return e;
        }
        uint32_t *counter = static_cast<uint32_t *>(store.counter);
        hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(64), 0, stream, counter);
        if (!workgroups) {
            int dev = 0, cus = 256;
            if (hipGetDevice(&dev) == hipSuccess)
                (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
            workgroups = uint32_t(cus > 0 ? cus : 256) * uint32_t(FPT_VOXEL_PERSISTENT_WAVES);
        }
        if (event_batch <= 0) event_batch = kVoxelEventBatch;
        const uint32_t slots = slot_count(hcfg.width, hcfg.height);
        const float *focus = static_cast<const float *>(store.focus);
        const bool b8 = store.storage == FPT_VOXEL_STORAGE_SPARSE_BRICKS8;
#define FPT_VX_LAUNCH(P, B8)                                                                                       \
    hipLaunchKernelGGL((voxel_persistent_kernel<P, B8>), dim3(workgroups), dim3(256), 0, stream, dcfg, v, focus,   \
                       accum, start, count, counter, slots, event_batch, dda)
        if (dda) {
            if (b8) FPT_VX_LAUNCH(true, true);
            else FPT_VX_LAUNCH(true, false);
        } else {
            if (b8) FPT_VX_LAUNCH(false, true);
            else FPT_VX_LAUNCH(false, false);
        }
#undef FPT_VX_LAUNCH
        return hipGetLastError();
    }
    dim3 grid((hcfg.width + 15u) / 16u, (hcfg.height + 15u) / 16u, 1u);
    if (voxel_needs_field(hcfg))
        hipLaunchKernelGGL(voxel_accumulate_kernel, grid, dim3(256), 0, stream, dcfg, v,
                           static_cast<const float *>(store.focus), accum, start, count, dda);
    else
        hipLaunchKernelGGL(voxel_accumulate_table_kernel, grid, dim3(256), 0, stream, dcfg, v,
                           static_cast<const float *>(store.focus), accum, start, count, dda);
    return hipGetLastError();
}

// voxelDiagnostic (Shaders.metal:7390-7442).  PATH_FINAL sums
// min(max(spp, 1), 64) voxel paths; the rest look at the un-jittered camera
// ray's first voxel hit.
__device__ f3 voxel_diagnostic(UV xy, const FptRenderConfig &cfg, const FptDiagnosticConfig &dg, const VoxelView &v,
                               float focus, Work &w) {
    const uint32_t mode = dg.mode;
    int unused = 0;
    DdaWork dw = {0u, 0u, 0u};
    if (mode == FPT_DIAGNOSTIC_PATH_FINAL) {
        const uint32_t samples = cfg.samples < 1u ? 1u : (cfg.samples > 64u ? 64u : cfg.samples);
        f3 c = mk(0.f, 0.f, 0.f);
        for (uint32_t s = 0; s < samples; ++s) c = c + render_voxel_path<true>(xy, s, focus, cfg, v, w, unused, dw);
        return post_process(c / float(samples), cfg);
    }
    const float focal = 1.0f / tanf(cfg.camera_fov / 2.0f * kPi / 180.0f);
    const f3 dir = rotate_camera_roll(normalize(mk(xy.x, xy.y, focal)), cfg.camera_yaw_pitch[0],
                                      cfg.camera_yaw_pitch[1], cfg.camera_roll);
    VHit h = trace_voxel(ld3(cfg.camera_position), dir, cfg, v, NoVisit());
    refine_leaf(h, ld3(cfg.camera_position), dir, cfg, w);
    if (mode == FPT_DIAGNOSTIC_HIT_MASK) return h.hit ? mk(1.f, 1.f, 1.f) : mk(0.f, 0.f, 0.f);
    if (mode == FPT_DIAGNOSTIC_DEPTH) {  // SM:6827-6832 on the surface distance (SM:7410)
        if (!h.hit) return mk(0.f, 0.f, 0.f);
        const float md = dg.max_distance > 0.001f ? dg.max_distance : cfg.render[4];
        const float k = 1.0f - clampf(h.surface / fmaxf(md, 0.001f), 0.f, 1.f);
        return mk(k, k, k);
    }
    if (mode == FPT_DIAGNOSTIC_SDF_PRIMARY_STEPS) {
        const float scale = float((cfg.voxel_resolution > 1u ? cfg.voxel_resolution : 1u) * 3u);
        const float k = clampf(float(h.steps) / fmaxf(scale, 1.0f), 0.f, 1.f);
        return mk(k, k, k);
    }
    if (!h.hit) return mk(0.f, 0.f, 0.f);
    f3 n = cfg.voxel_normal_mode == 1u ? normal_scene(h.cell_center, cfg, w)
                                        : (cfg.voxel_normal_mode == 2u ? normal_scene(h.position, cfg, w) : h.normal);
    if (dot(n, dir) > 0.0f) n = -n;
    if (mode == FPT_DIAGNOSTIC_DIFFUSE_NORMAL) {
        const f3 fwd = rotate_camera_roll(mk(0.f, 0.f, 1.f), cfg.camera_yaw_pitch[0], cfg.camera_yaw_pitch[1],
                                          cfg.camera_roll);
        const float k = 0.12f + 0.88f * fmaxf(dot(n, -fwd), 0.0f);
        return mk(k, k, k);
    }
    f3 on = cfg.voxel_normal_mode == 2u ? h.normal : n;
    if (dot(on, dir) > 0.0f) on = -on;
    const Material mat = hit_material(h, cfg, v, w);
    if (mode == FPT_DIAGNOSTIC_MATERIAL)
        return mk(clampf(mat.rgb.x, 0.f, 1.f), clampf(mat.rgb.y, 0.f, 1.f), clampf(mat.rgb.z, 0.f, 1.f));
    if (mode == FPT_DIAGNOSTIC_PATH_DIRECT) {
        const f3 c = cfg.sun[0] == 1.0f ? voxel_sun(h.position, xy, 0.0f, mat, n, on, cfg, v, unused, dw)
                                        : mk(0.f, 0.f, 0.f);
        return post_process(vmin(vmax0(c), 8.0f), cfg);
    }
    if (mode == FPT_DIAGNOSTIC_PATH_ENVIRONMENT) return post_process(vmin(vmax0(environment(dir, cfg)), 8.0f), cfg);
    if (mode == FPT_DIAGNOSTIC_PATH_THROUGHPUT) {
        const f3 next = normalize(vmix(reflect(dir, n), random_vector(n, xy, 0.0f), mat.roughness));
        const float al = clampf(dot(next, n) * 0.5f + 0.5f, 0.f, 1.f);
        const f3 c = mat.rgb * al;
        return mk(clampf(c.x, 0.f, 1.f), clampf(c.y, 0.f, 1.f), clampf(c.z, 0.f, 1.f));
    }
    return mk(clampf(n.x * 0.5f + 0.5f, 0.f, 1.f), clampf(n.y * 0.5f + 0.5f, 0.f, 1.f),
              clampf(n.z * 0.5f + 0.5f, 0.f, 1.f));
}

__global__ __launch_bounds__(256) void voxel_diagnostic_kernel(const FptRenderConfig *__restrict__ cfgp, VoxelView v,
                                                               const float *__restrict__ focus_p,
                                                               FptDiagnosticConfig dg, uchar4 *__restrict__ out) {
    const FptRenderConfig &cfg = *cfgp;
    const uint32_t t = threadIdx.x, wave = t >> 6, lane = t & 63u;
    const uint32_t x = blockIdx.x * 16u + (wave & 1u) * 8u + (lane & 7u);
    const uint32_t y = blockIdx.y * 16u + (wave >> 1) * 8u + (lane >> 3);
    if (x >= cfg.width || y >= cfg.height) return;
    Work w = {0, 0};
    const f3 c = voxel_diagnostic(screen_uv(x, y, cfg.width, cfg.height), cfg, dg, v, *focus_p, w);
    out[size_t(cfg.height - 1u - y) * cfg.width + x] =
        make_uchar4((unsigned char)(clampf(c.x, 0.f, 1.f) * 255.0f), (unsigned char)(clampf(c.y, 0.f, 1.f) * 255.0f),
                    (unsigned char)(clampf(c.z, 0.f, 1.f) * 255.0f), 255);
}

hipError_t launch_voxel_diagnostic(const FptRenderConfig *dcfg, const FptRenderConfig &hcfg, VoxelStore &store,
                                   const FptDiagnosticConfig &dg, uchar4 *out, hipStream_t stream) {
    VoxelView v = view_of(store);
    if (!store.focus) {
        hipError_t e = hipMalloc(&store.focus, sizeof(float));
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(voxel_focus_kernel, dim3(1), dim3(64), 0, stream, dcfg, v, static_cast<float *>(store.focus));
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    dim3 grid((hcfg.width + 15u) / 16u, (hcfg.height + 15u) / 16u, 1u);
    hipLaunchKernelGGL(voxel_diagnostic_kernel, grid, dim3(256), 0, stream, dcfg, v,
                       static_cast<const float *>(store.focus), dg, out);
    return hipGetLastError();
}

hipError_t launch_voxel_trace_cells(const FptRenderConfig *dcfg, const VoxelStore &store, const float *rays, size_t n,
                                    uint32_t cap, int *cells, uint32_t *counts, float *hits, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(voxel_trace_cells_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream, dcfg,
                       view_of(store), rays, n, cap, cells, counts, hits);
    return hipGetLastError();
}

}  // namespace fpt

namespace {
// The typed programs of the reference's bounds contract
// (metal_bridge.mm:3748-3790): every config a zero-initialised typed-program
// scene with surface band 0.5.
void contract_programs(FptRenderConfig *pc) {
    std::memset(pc, 0, 9 * sizeof(FptRenderConfig));
    for (int i = 0; i < 9; ++i) {
        pc[i].sdf_id = FPT_SDF_PROGRAM;
        pc[i].voxel_surface_band = 0.5f;
    }
    auto add = [](FptRenderConfig &c, uint32_t op, uint32_t flags, float x, float y, float z, float w) {
        FptSdfInstruction &in = c.sdf_program[c.sdf_program_count++];
        in.opcode = op;
        in.flags = flags;
        in.data[0] = x; in.data[1] = y; in.data[2] = z; in.data[3] = w;
    };
    add(pc[0], FPT_SDF_OP_SPHERE, 0, 1.0f, 0, 0, 0);
    add(pc[1], FPT_SDF_OP_BOX, 0, 0.8f, 0.6f, 0.4f, 0);
    add(pc[2], FPT_SDF_OP_PLANE, 0, 1.0f, 2.0f, -0.5f, 0.2f);
    add(pc[3], FPT_SDF_OP_REPEAT, 0, 4.0f, 4.0f, 4.0f, 0);
    add(pc[3], FPT_SDF_OP_SPHERE, 0, 1.0f, 0, 0, 0);
    add(pc[4], FPT_SDF_OP_ABS, 0, 0, 0, 0, 0);
    add(pc[4], FPT_SDF_OP_ROTATE_X, 0, 0.31f, 0, 0, 0);
    add(pc[4], FPT_SDF_OP_ROTATE_Y, 0, -0.47f, 0, 0, 0);
    add(pc[4], FPT_SDF_OP_ROTATE_Z, 0, 0.19f, 0, 0, 0);
    add(pc[4], FPT_SDF_OP_SCALE, 0, 1.25f, 0, 0, 0);
    add(pc[4], FPT_SDF_OP_TRANSLATE, 0, 0.2f, -0.1f, 0.3f, 0);
    add(pc[4], FPT_SDF_OP_BOX, 0, 1.0f, 0.8f, 0.6f, 0);
    add(pc[4], FPT_SDF_OP_SPHERE, 2, 0.35f, 0, 0, 0);
    add(pc[5], FPT_SDF_OP_SPHERE, 0, 1.0f, 0, 0, 0);
    add(pc[5], FPT_SDF_OP_TRANSLATE, 0, 0.5f, 0, 0, 0);
    add(pc[5], FPT_SDF_OP_BOX, 1, 0.7f, 0.7f, 0.7f, 0);
    add(pc[6], FPT_SDF_OP_SPHERE, 0, 0.7f, 0, 0, 0);
    add(pc[6], FPT_SDF_OP_TRANSLATE, 0, 0.9f, 0, 0, 0);
    add(pc[6], FPT_SDF_OP_SPHERE, 0, 0.7f, 0, 0, 0);
    add(pc[7], FPT_SDF_OP_SORT_DESC, 0, 0, 0, 0, 0);
    add(pc[7], FPT_SDF_OP_SPHERE, 0, 1.0f, 0, 0, 0);
    add(pc[8], FPT_SDF_OP_SORT_DESC, 0, 0, 0, 0, 0);
    add(pc[8], FPT_SDF_OP_BOX, 0, 0.9f, 0.55f, 0.3f, 0);
}
}  // namespace

// The reference's direct-build parity check (metal_bridge.mm:3836-3923): the
// contract's unit sphere on a 32^3 grid over [-2, 2]^3, 4^3 bricks, interval
// coverage, built by staging and by the direct build with certified brick
// rejection.  Interval coverage keeps no cell in a brick whose bounds exclude
// the surface, so the two builds must hold the same bricks and cells, cell
// for cell through their page tables.
static int direct_build_parity(const FptRenderConfig &sphere, char *error, size_t error_len) {
    FptRenderConfig c = sphere;
    c.voxel_resolution = 32u;
    c.voxel_storage = FPT_VOXEL_STORAGE_SPARSE_BRICKS;
    c.voxel_coverage_mode = FPT_VOXEL_COVERAGE_INTERVAL;
    c.voxel_brick_rejection = 1u;
    for (int a = 0; a < 3; ++a) {
        c.voxel_bounds_min[a] = -2.0f;
        c.voxel_bounds_max[a] = 2.0f;
    }
    FptRenderConfig cfgs[2] = {c, c};
    cfgs[0].voxel_build_mode = FPT_VOXEL_BUILD_STAGING;
    cfgs[1].voxel_build_mode = FPT_VOXEL_BUILD_DIRECT;
    struct Tables {
        std::vector<uint32_t> pages;
        std::vector<fpt::VoxelRecord> cells;
        uint32_t bricks = 0, rejected = 0;
        uint64_t active = 0;
    } t[2];
    FptRenderConfig *d = nullptr;
    if (hipMalloc(&d, sizeof(FptRenderConfig)) != hipSuccess) {
        if (error && error_len) std::snprintf(error, error_len, "failed to allocate direct-build parity resources");
        return 1;
    }
    for (int k = 0; k < 2; ++k) {
        fpt::VoxelStore vs;
        hipError_t e = hipMemcpy(d, &cfgs[k], sizeof(FptRenderConfig), hipMemcpyHostToDevice);
        if (e != hipSuccess || fpt::voxel_build(cfgs[k], d, vs, nullptr, error, error_len)) {
            fpt::voxel_free(vs);
            (void)hipFree(d);
            if (e != hipSuccess && error && error_len)
                std::snprintf(error, error_len, "direct-build parity upload failed: %s", hipGetErrorString(e));
            return 1;
        }
        t[k].pages.resize(size_t(vs.bricks_per_axis) * vs.bricks_per_axis * vs.bricks_per_axis);
        t[k].cells.resize(size_t(vs.active_bricks) * 64u);
        t[k].bricks = vs.active_bricks;
        t[k].rejected = vs.rejected_bricks;
        t[k].active = vs.active_cells;
        e = hipMemcpy(t[k].pages.data(), vs.page_table, t[k].pages.size() * 4u, hipMemcpyDeviceToHost);
        if (e == hipSuccess && !t[k].cells.empty())
            e = hipMemcpy(t[k].cells.data(), vs.pool, t[k].cells.size() * sizeof(fpt::VoxelRecord),
                          hipMemcpyDeviceToHost);
        fpt::voxel_free(vs);
        if (e != hipSuccess) {
            (void)hipFree(d);
            if (error && error_len)
                std::snprintf(error, error_len, "direct-build parity readback failed: %s", hipGetErrorString(e));
            return 1;
        }
    }
    (void)hipFree(d);
    if (t[0].bricks != t[1].bricks || t[0].active != t[1].active || t[1].rejected == 0u) {
        if (error && error_len) std::snprintf(error, error_len, "direct-build occupancy totals differ from staging");
        return 1;
    }
    const fpt::VoxelRecord empty = {0u, 0u, 0.0f};
    for (uint32_t z = 0; z < 32u; ++z)
        for (uint32_t y = 0; y < 32u; ++y)
            for (uint32_t x = 0; x < 32u; ++x) {
                const uint32_t page_index = x / 4u + (y / 4u) * 8u + (z / 4u) * 64u;
                const uint32_t local = x % 4u + (y % 4u) * 4u + (z % 4u) * 16u;
                const fpt::VoxelRecord *cell[2];
                for (int k = 0; k < 2; ++k) {
                    const uint32_t page = t[k].pages[page_index];
                    cell[k] = page == 0u ? &empty : &t[k].cells[size_t(page - 1u) * 64u + local];
                }
                if (std::memcmp(cell[0], cell[1], sizeof(fpt::VoxelRecord)) != 0) {
                    if (error && error_len)
                        std::snprintf(error, error_len, "direct-build cell mismatch at (%u,%u,%u)", x, y, z);
                    return 1;
                }
            }
    return 0;
}

// Runs the known-answer checks of the reference's contract (Shaders.metal:
// 5156-5386) that this library implements: bits 0-29.  Bits 30 (Cage-fractal
// bounds) and 31 (derivative bounds) belong to features this library does not
// build.
extern "C" int fpt_test_voxel_dda(const char *metallib_path, char *error, size_t error_len) {
    (void)metallib_path;
    std::vector<FptRenderConfig> pc(9);
    contract_programs(pc.data());
    uint32_t *d = nullptr;
    FptRenderConfig *dpc = nullptr;
    hipError_t e = hipMalloc(&d, 4);
    if (e == hipSuccess) e = hipMalloc(&dpc, pc.size() * sizeof(FptRenderConfig));
    if (e == hipSuccess) e = hipMemcpy(dpc, pc.data(), pc.size() * sizeof(FptRenderConfig), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(fpt::voxel_dda_contract_kernel, dim3(1), dim3(1), 0, 0, d, dpc);
        e = hipGetLastError();
    }
    uint32_t f = 0xffffffffu;
    if (e == hipSuccess) e = hipMemcpy(&f, d, 4, hipMemcpyDeviceToHost);
    if (d) (void)hipFree(d);
    if (dpc) (void)hipFree(dpc);
    if (e != hipSuccess) {
        if (error && error_len) std::snprintf(error, error_len, "voxel DDA contract could not run: %s", hipGetErrorString(e));
        return 1;
    }
    if (f != 0u) {
        if (error && error_len) std::snprintf(error, error_len, "voxel DDA contract failures: 0x%08x", f);
        return 1;
    }
    return direct_build_parity(pc[0], error, error_len);
}
