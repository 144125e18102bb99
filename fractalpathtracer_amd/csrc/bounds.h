// Certified distance bounds of a typed program over an axis-aligned cell and
// the voxel-cell classification built on them (Shaders.metal:1065-1273,
// 4196-4244).
//
// program_bounds() pushes the cell (centre, half extent) through the
// program's transforms as an interval box: abs folds it onto the positive
// octant, rotations grow it to the rotated box's hull, repeat stays exact
// while the box does not straddle a period seam, sort takes the hull of the
// sorted corners; each primitive then gives [min, max] of its distance over
// the box, combined with the CSG rule of its flags.  Fractal leaves and the
// box fold have no interval form here, so a program that contains one has no
// certified bounds (as any opcode the reference does not bound).  The result
// is widened outward by an explicit rounding allowance and one ulp.
#pragma once

#include "fpt_device.h"

namespace fpt {

struct ProgramBounds {
    float minimum, maximum, lipschitz;
    bool certified;
};

enum : uint32_t { CELL_OUTSIDE = 0u, CELL_INSIDE = 1u, CELL_SURFACE = 2u, CELL_UNKNOWN = 3u };

struct CellClass {
    float interval_min, interval_max;
    bool certified, legacy, lipschitz, interval;
    uint32_t cell_class;
};

__device__ __forceinline__ ProgramBounds unknown_bounds() {
    return {-__builtin_inff(), __builtin_inff(), __builtin_inff(), false};
}

// Next float towards -inf / +inf (zero steps to the smallest subnormal).
__device__ __forceinline__ float round_down(float v) {
    if (!isfinite(v)) return v;
    if (v == 0.0f) return -__uint_as_float(1u);
    const uint32_t b = __float_as_uint(v);
    return __uint_as_float(v < 0.0f ? b + 1u : b - 1u);
}
__device__ __forceinline__ float round_up(float v) {
    if (!isfinite(v)) return v;
    if (v == 0.0f) return __uint_as_float(1u);
    const uint32_t b = __float_as_uint(v);
    return __uint_as_float(v < 0.0f ? b - 1u : b + 1u);
}

__device__ __forceinline__ ProgramBounds combine_bounds(ProgramBounds a, ProgramBounds b, uint32_t mode) {
    if (!a.certified || !b.certified) return unknown_bounds();
    ProgramBounds r;
    if (mode == 1u) {
        r.minimum = fmaxf(a.minimum, b.minimum);
        r.maximum = fmaxf(a.maximum, b.maximum);
    } else if (mode == 2u) {
        r.minimum = fmaxf(a.minimum, -b.maximum);
        r.maximum = fmaxf(a.maximum, -b.minimum);
    } else {
        r.minimum = fminf(a.minimum, b.minimum);
        r.maximum = fminf(a.maximum, b.maximum);
    }
    r.lipschitz = fmaxf(a.lipschitz, b.lipschitz);
    r.certified = true;
    return r;
}

// Rotate the box centre and grow the half extent to the hull of the rotated
// box in the rotation plane (a, b).
__device__ __forceinline__ void rotate_box(float &ca, float &cb, float &ha, float &hb, float angle) {
    float s, c;
    sincosf(angle, &s, &c);
    const float na = c * ca - s * cb;
    cb = s * ca + c * cb;
    ca = na;
    const float ac = fabsf(c), as = fabsf(s);
    const float ea = ha, eb = hb;
    ha = ac * ea + as * eb;
    hb = as * ea + ac * eb;
}

__device__ __forceinline__ f3 sort_desc3(f3 v) {  // SM:1127-1132
    if (v.x < v.z) { float t = v.x; v.x = v.z; v.z = t; }
    if (v.y < v.z) { float t = v.y; v.y = v.z; v.z = t; }
    if (v.x < v.y) { float t = v.x; v.x = v.y; v.y = t; }
    return v;
}

__device__ ProgramBounds program_bounds(f3 center, f3 half, const FptRenderConfig &cfg) {  // SM:1134-1273
    f3 h = vabs(half);
    float ds = 1.0f;
    ProgramBounds acc = unknown_bounds();
    bool has = false, certified = true;
    const uint32_t count = cfg.sdf_program_count < uint32_t(FPT_SDF_PROGRAM_MAX_OPS) ? cfg.sdf_program_count
                                                                                     : uint32_t(FPT_SDF_PROGRAM_MAX_OPS);
    for (uint32_t i = 0; i < count; ++i) {
        const FptSdfInstruction &op = cfg.sdf_program[i];
        const float *d = op.data;
        ProgramBounds prim;
        bool leaf = false;
        switch (op.opcode) {
            case FPT_SDF_OP_ABS: {
                const f3 lo = center - h, hi = center + h;
                auto one = [](float l, float u, float &c, float &e) {
                    const float amin = (l <= 0.0f && u >= 0.0f) ? 0.0f : fminf(fabsf(l), fabsf(u));
                    const float amax = fmaxf(fabsf(l), fabsf(u));
                    c = (amin + amax) * 0.5f;
                    e = (amax - amin) * 0.5f;
                };
                one(lo.x, hi.x, center.x, h.x);
                one(lo.y, hi.y, center.y, h.y);
                one(lo.z, hi.z, center.z, h.z);
                break;
            }
            case FPT_SDF_OP_TRANSLATE: center = center - ld3(d); break;
            case FPT_SDF_OP_SCALE: {
                const float s = fabsf(d[0]) > 1.0e-6f ? d[0] : 1.0f;
                center = center * s;
                h = h * fabsf(s);
                ds *= fabsf(s);
                break;
            }
            case FPT_SDF_OP_ROTATE_X: rotate_box(center.y, center.z, h.y, h.z, d[0]); break;
            case FPT_SDF_OP_ROTATE_Y: rotate_box(center.x, center.z, h.x, h.z, d[0]); break;
            case FPT_SDF_OP_ROTATE_Z: rotate_box(center.x, center.y, h.x, h.y, d[0]); break;
            case FPT_SDF_OP_REPEAT: {
                const f3 per = mk(fmaxf(fabsf(d[0]), 1.0e-5f), fmaxf(fabsf(d[1]), 1.0e-5f), fmaxf(fabsf(d[2]), 1.0e-5f));
                const f3 lo = center - h, hi = center + h;
                if (floorf(lo.x / per.x + 0.5f) != floorf(hi.x / per.x + 0.5f) ||
                    floorf(lo.y / per.y + 0.5f) != floorf(hi.y / per.y + 0.5f) ||
                    floorf(lo.z / per.z + 0.5f) != floorf(hi.z / per.z + 0.5f))
                    certified = false;
                center.x -= per.x * floorf(center.x / per.x + 0.5f);
                center.y -= per.y * floorf(center.y / per.y + 0.5f);
                center.z -= per.z * floorf(center.z / per.z + 0.5f);
                break;
            }
            case FPT_SDF_OP_SORT_DESC: {
                const f3 lo = sort_desc3(center - h), hi = sort_desc3(center + h);
                center = (lo + hi) * 0.5f;
                h = vmax0((hi - lo) * 0.5f);
                break;
            }
            case FPT_SDF_OP_SPHERE: {
                const float div = fmaxf(ds, 1.0e-6f);
                prim.minimum = (length(vmax0(vabs(center) - h)) - d[0]) / div;
                prim.maximum = (length(vabs(center) + h) - d[0]) / div;
                prim.lipschitz = 1.0f;
                prim.certified = certified;
                leaf = true;
                break;
            }
            case FPT_SDF_OP_BOX: {
                const float div = fmaxf(ds, 1.0e-6f);
                const f3 q = vabs(center) - vabs(ld3(d));
                const float cd = (fminf(fmaxf(q.x, fmaxf(q.y, q.z)), 0.0f) + length(vmax0(q))) / div;
                const float r = length(h) / div;
                prim.minimum = cd - r;
                prim.maximum = cd + r;
                prim.lipschitz = 1.0f;
                prim.certified = certified;
                leaf = true;
                break;
            }
            case FPT_SDF_OP_PLANE: {
                const float nl = length(ld3(d));
                const float div = fmaxf(ds, 1.0e-6f);
                if (nl <= 1.0e-8f) {
                    prim = unknown_bounds();
                } else {
                    const f3 n = ld3(d) / nl;
                    const float mid = (dot(center, n) + d[3]) / div;
                    const float r = dot(h, vabs(n)) / div;
                    prim.minimum = mid - r;
                    prim.maximum = mid + r;
                    prim.lipschitz = 1.0f;
                    prim.certified = certified;
                }
                leaf = true;
                break;
            }
            case FPT_SDF_OP_ORBIT_ADD:
            case FPT_SDF_OP_UNION:
            case FPT_SDF_OP_INTERSECTION:
            case FPT_SDF_OP_SUBTRACT:
                break;
            default:  // fractal leaves, box fold, unknown opcodes
                certified = false;
                break;
        }
        if (leaf) {
            acc = has ? combine_bounds(acc, prim, op.flags) : prim;
            has = true;
        }
    }
    if (!has || !certified || !acc.certified || !isfinite(acc.minimum) || !isfinite(acc.maximum)) return unknown_bounds();
    // Outward allowance for the float evaluation of the interval ops: grows
    // with program length and coordinate magnitude, floored at 2e-5 (1 + |d|).
    const float eps = 1.1920929e-7f;
    const float mag = fmaxf(fabsf(acc.minimum), fabsf(acc.maximum));
    const float scale = 1.0f + mag + length(center) + length(h) + fabsf(ds);
    const float allowance = fmaxf(2.0e-5f * (1.0f + mag), scale * float(count + 1u) * 8.0f * eps);
    acc.minimum = round_down(acc.minimum - allowance);
    acc.maximum = round_up(acc.maximum + allowance);
    return acc;
}

// classifyAccelerationCell (SM:4196-4244): legacy occupancy from the centre
// distance d = distanceSdf(center), and for typed programs the Lipschitz and interval occupancies and
// the cell class from certified bounds.
__device__ CellClass classify_cell(f3 center, f3 half, float d, const FptRenderConfig &cfg) {
    CellClass r;
    const float radius = length(half);
    const bool fin = isfinite(d);
    r.interval_min = fin ? d - radius : -__builtin_inff();
    r.interval_max = fin ? d + radius : __builtin_inff();
    r.certified = false;
    bool occ = fin && fabsf(d) <= radius * fmaxf(cfg.voxel_surface_band, 0.25f);
    const bool fill = cfg.voxel_fill_interior != 0u && fin && d < 0.0f;
    occ = occ || fill;
    r.legacy = r.lipschitz = r.interval = occ;
    r.cell_class = CELL_UNKNOWN;
    if (cfg.sdf_id == FPT_SDF_PROGRAM && cfg.sdf_program_count > 0u) {
        const ProgramBounds b = program_bounds(center, half, cfg);
        if (b.certified) {
            r.interval_min = b.minimum;
            r.interval_max = b.maximum;
            r.certified = true;
            r.lipschitz = (fin && fabsf(d) <= b.lipschitz * radius) || fill;
            r.interval = (b.minimum <= 0.0f && b.maximum >= 0.0f) || fill;
            r.cell_class = b.minimum > 0.0f ? CELL_OUTSIDE : (b.maximum < 0.0f ? CELL_INSIDE : CELL_SURFACE);
        }
    }
    return r;
}

}  // namespace fpt
