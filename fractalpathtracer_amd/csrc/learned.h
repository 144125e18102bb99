// Learned far-field step proposer: a small per-scene MLP that estimates the
// scene distance, evaluated wave-collectively with BF16 MFMA on gfx950.
// There is no counterpart in the reference; the exact distance estimator
// (Shaders.metal:3485-3615) stays the authority near the surface.
//
// Network (logical form; float parameter layout in fpt_abi.h):
//   v        = p / box_half                      (training box -> [-1, 1]^3)
//   F[128]   feature vector, index k = 32*a + 8*q + 2*i + c for axis a < 3,
//            quarter q < 4, i < 4, c = 0 (sin) / 1 (cos):
//              F = sin|cos(pi * f * v_a),  f = 2^((4q + i) / 2)
//            (16 frequencies per axis), then F[96..98] = v, F[99..127] = 0
//   h1 = silu(W1 F + b1)          64 x 128
//   h2 = silu(W2 h1 + b2)         64 x 64
//   h3 = silu(W3 h2 + b3)
//   h4 = silu(W4 h3 + b4)
//   out = w5 . h4 + b5
// Weights and the inputs of W1..W4 (F, h1, h2, h3) are rounded to BF16
// (round to nearest even); products accumulate in FP32; h4 and the output
// layer stay FP32.
//
// MFMA mapping (v_mfma_f32_16x16x32_bf16, D = A B + C, 16 x 16 x 32): the
// layers are computed transposed, H^T = W X^T, so the weight rows (units) are
// the M dimension and the wave's 64 points the N dimension, one 16-point
// column tile `nt` at a time.  Lane l holds column (point) 16*nt + (l & 15)
// and rows 4*(l >> 4) + r of each 16-row accumulator tile: exactly the units
// the next layer's B operand needs in that lane, once the hidden-layer
// weights are stored with their K index permuted to match (see the packing in
// learned.hip).  The whole network therefore runs in registers: no LDS, no
// cross-lane traffic except the broadcast of the point coordinates and two
// xor-shuffles for the final dot product.
#pragma once

#ifndef __HIPCC_RTC__  // the runtime compiler provides these itself
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

#include "fpt_device.h"

namespace fpt {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kProposerInputs = 128;
constexpr int kProposerHidden = 64;
constexpr int kProposerFrags = 16 + 3 * 8;  // A fragments: layer 1 (4 x 4), hidden (3 x 4 x 2)

struct ProposerDevice {
    const bf16x8 *frags;  // kProposerFrags x 64 lanes x 8 bf16, fragment-major
    const float *bias;    // b1 | b2 | b3 | b4 (4 x 64)
    const float *w_out;   // w5 (64)
    float b_out;          // b5
    float inv_half;       // 1 / box_half
    float box_half;       // training box half extent (box centred at the origin)
    float switch_distance;  // proposal at or below this: hand over to the exact DE
    float resume_distance;  // exact |d| above this: proposals allowed again
    float step_scale;       // far step = step_scale * proposal
};

__device__ __forceinline__ float proposer_silu(float x) { return x * __builtin_amdgcn_rcpf(1.0f + __expf(-x)); }

// sin / cos of (2 pi rev) with the hardware sine (input in revolutions,
// reduced to [0, 1) first).
__device__ __forceinline__ void sincos_rev(float rev, float &s, float &c) {
    const float r = __builtin_amdgcn_fractf(rev);
    s = __builtin_amdgcn_sinf(r);
    c = __builtin_amdgcn_cosf(r);
}

// The 8 features of quarter q for one axis value v: frequencies
// f = 2^(2q) * {1, sqrt2, 2, 2 sqrt2}; pi f v = 2 pi (f v / 2) revolutions.
__device__ __forceinline__ bf16x8 axis_features(float v, uint32_t q) {
    const float base = 0.5f * float(1u << (2u * q)) * v;
    float s0, c0, s1, c1, s2, c2, s3, c3;
    sincos_rev(base, s0, c0);
    sincos_rev(base * 1.41421356237309515f, s1, c1);
    sincos_rev(base * 2.0f, s2, c2);
    sincos_rev(base * 2.82842712474619030f, s3, c3);
    return bf16x8{(__bf16)s0, (__bf16)c0, (__bf16)s1, (__bf16)c1, (__bf16)s2, (__bf16)c2, (__bf16)s3, (__bf16)c3};
}

// Bias + SiLU of two accumulator tiles (units 16*m + 4*(l>>4) + r for
// m = m0, m0 + 1), packed as the next layer's B operand.
__device__ __forceinline__ bf16x8 silu_pack(f32x4 c0, f32x4 c1, const float *b0, const float *b1) {
    return bf16x8{(__bf16)proposer_silu(c0[0] + b0[0]), (__bf16)proposer_silu(c0[1] + b0[1]),
                  (__bf16)proposer_silu(c0[2] + b0[2]), (__bf16)proposer_silu(c0[3] + b0[3]),
                  (__bf16)proposer_silu(c1[0] + b1[0]), (__bf16)proposer_silu(c1[1] + b1[1]),
                  (__bf16)proposer_silu(c1[2] + b1[2]), (__bf16)proposer_silu(c1[3] + b1[3])};
}

// Distance proposal for this lane's point `pos`.  Wave-collective: all 64
// lanes of the wave must execute it with the same ProposerDevice (lanes with
// nothing to ask pass any finite point and ignore the result).
__device__ __forceinline__ float proposer_eval(const ProposerDevice &P, f3 pos) {
    const uint32_t l = threadIdx.x & 63u, lr = l & 15u, lq = l >> 4;
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    float result = 0.f;
#pragma unroll 1
    for (uint32_t nt = 0; nt < 4; ++nt) {
        const int src = int(16u * nt + lr);
        const float vx = __shfl(pos.x, src, 64) * P.inv_half;
        const float vy = __shfl(pos.y, src, 64) * P.inv_half;
        const float vz = __shfl(pos.z, src, 64) * P.inv_half;
        bf16x8 b[4];
        b[0] = axis_features(vx, lq);
        b[1] = axis_features(vy, lq);
        b[2] = axis_features(vz, lq);
        const __bf16 bz = (__bf16)0.0f;
        b[3] = lq == 0u ? bf16x8{(__bf16)vx, (__bf16)vy, (__bf16)vz, bz, bz, bz, bz, bz}
                        : bf16x8{bz, bz, bz, bz, bz, bz, bz, bz};
        // Layer 1: 4 row tiles x 4 K steps.
        f32x4 c[4];
        const bf16x8 *fr = P.frags + l;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            c[mt] = zero;
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)
                c[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fr[(mt * 4 + ks) * 64], b[ks], c[mt], 0, 0, 0);
        }
        // Hidden layers 2-4: 4 row tiles x 2 K steps.
#pragma unroll
        for (int h = 0; h < 3; ++h) {
            const float *bias = P.bias + 64 * h + 4 * lq;
            bf16x8 a0 = silu_pack(c[0], c[1], bias, bias + 16);
            bf16x8 a1 = silu_pack(c[2], c[3], bias + 32, bias + 48);
            const bf16x8 *fh = P.frags + (16 + 8 * h) * 64 + l;
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                c[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fh[(mt * 2 + 0) * 64], a0, zero, 0, 0, 0);
                c[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fh[(mt * 2 + 1) * 64], a1, c[mt], 0, 0, 0);
            }
        }
        // Output: this lane's 16 units of h4, then the sum over the four
        // lanes that share the column.
        const float *b4 = P.bias + 192 + 4 * lq;
        const float *w5 = P.w_out + 4 * lq;
        float part = 0.f;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int r = 0; r < 4; ++r) part += w5[16 * mt + r] * proposer_silu(c[mt][r] + b4[16 * mt + r]);
        part += __shfl_xor(part, 16, 64);
        part += __shfl_xor(part, 32, 64);
        if (nt == lq) result = part;  // column lr of tile lq is this lane's own point
    }
    return result + P.b_out;
}

__device__ __forceinline__ bool proposer_in_box(const ProposerDevice &P, f3 p) {
    return fabsf(p.x) < P.box_half && fabsf(p.y) < P.box_half && fabsf(p.z) < P.box_half;
}

#ifndef __HIPCC_RTC__
// Host side (learned.hip).
struct ProposerHost {
    void *blob = nullptr;  // device: frags | bias | w_out
    ProposerDevice dev{};
    bool enabled = false;
};
// Packs the logical float parameters (FPT_PROPOSER_PARAM_COUNT) into
// `host.blob` on the device; returns a message on failure.
const char *proposer_upload(ProposerHost &host, const float *params, size_t count, float box_half,
                            float switch_distance, float resume_distance, float step_scale);
void proposer_release(ProposerHost &host);
hipError_t launch_proposer_eval(const ProposerDevice &P, const float4 *pts, size_t n, float *out,
                                hipStream_t stream);
hipError_t launch_proposer_samples(const FptRenderConfig *dcfg, float box_half, float band, uint32_t count,
                                   uint64_t seed, uint32_t max_attempts, float4 *out, uint32_t *failures,
                                   hipStream_t stream);

#endif

}  // namespace fpt
