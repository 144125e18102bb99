// Learned step proposer: host-side weight packing, the standalone evaluation
// kernel (parity hook) and the training-sample generator.  The network and its
// MFMA mapping are described in learned.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "fpt_device.h"
#include "learned.h"

namespace fpt {

namespace {

// Logical parameter offsets (floats), matching FPT_PROPOSER_PARAM_COUNT.
constexpr size_t kW1 = 0;                           // [64][128]
constexpr size_t kB1 = kW1 + 64 * 128;              // [64]
constexpr size_t kWh = kB1 + 64;                    // 3 x ([64][64] then [64] bias)
constexpr size_t kHiddenStride = 64 * 64 + 64;
constexpr size_t kW5 = kWh + 3 * kHiddenStride;     // [64]
constexpr size_t kB5 = kW5 + 64;                    // scalar
constexpr size_t kParamCount = kB5 + 1;             // 20801

// float -> bf16 bits, round to nearest even (finite inputs; NaN kept quiet).
uint16_t bf16_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return uint16_t((u >> 16) | 0x0040u);
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return uint16_t(u >> 16);
}

}  // namespace

const char *proposer_upload(ProposerHost &host, const float *params, size_t count, float box_half,
                            float switch_distance, float resume_distance, float step_scale) {
    if (!params || count != kParamCount) return "proposer parameter array must hold FPT_PROPOSER_PARAM_COUNT floats";
    for (size_t i = 0; i < count; ++i)
        if (!std::isfinite(params[i])) return "proposer parameters must be finite";
    if (!(box_half > 0.0f) || !std::isfinite(box_half)) return "proposer box half extent must be positive";
    if (!(step_scale > 0.0f && step_scale <= 1.0f)) return "proposer step scale must be in (0, 1]";
    if (!(switch_distance > 0.0f) || !(resume_distance >= switch_distance))
        return "proposer switch distance must be positive and not above the resume distance";

    // Fragment-major A operands: fragment f, lane l, slot j at
    // (f * 64 + l) * 8 + j.  Lane l covers row 16*mt + (l & 15) and the K
    // slots 8*(l >> 4) + j of its 32-wide K step.
    std::vector<uint16_t> frag(size_t(kProposerFrags) * 64 * 8);
    for (int mt = 0; mt < 4; ++mt)
        for (int ks = 0; ks < 4; ++ks)
            for (int l = 0; l < 64; ++l)
                for (int j = 0; j < 8; ++j) {
                    const int u = 16 * mt + (l & 15), k = 32 * ks + 8 * (l >> 4) + j;
                    frag[(size_t((mt * 4 + ks) * 64 + l)) * 8 + j] = bf16_bits(params[kW1 + size_t(u) * 128 + k]);
                }
    // Hidden layers: the B operand slot (ks, l, j) holds unit
    // 16*(2 ks + (j >> 2)) + 4*(l >> 4) + (j & 3) of the previous layer (the
    // accumulator layout), so the weight K index is permuted the same way.
    for (int h = 0; h < 3; ++h) {
        const float *W = params + kWh + size_t(h) * kHiddenStride;
        for (int mt = 0; mt < 4; ++mt)
            for (int ks = 0; ks < 2; ++ks)
                for (int l = 0; l < 64; ++l)
                    for (int j = 0; j < 8; ++j) {
                        const int u = 16 * mt + (l & 15);
                        const int k = 16 * (2 * ks + (j >> 2)) + 4 * (l >> 4) + (j & 3);
                        frag[(size_t((16 + 8 * h + mt * 2 + ks) * 64 + l)) * 8 + j] = bf16_bits(W[size_t(u) * 64 + k]);
                    }
    }
    std::vector<float> tail(4 * 64 + 64);
    std::memcpy(tail.data(), params + kB1, 64 * sizeof(float));
    for (int h = 0; h < 3; ++h)
        std::memcpy(tail.data() + 64 * (h + 1), params + kWh + size_t(h) * kHiddenStride + 64 * 64, 64 * sizeof(float));
    std::memcpy(tail.data() + 256, params + kW5, 64 * sizeof(float));

    const size_t frag_bytes = frag.size() * sizeof(uint16_t);
    const size_t bytes = frag_bytes + tail.size() * sizeof(float);
    proposer_release(host);
    if (hipMalloc(&host.blob, bytes) != hipSuccess) {
        host.blob = nullptr;
        return "proposer weight allocation failed";
    }
    if (hipMemcpy(host.blob, frag.data(), frag_bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(static_cast<char *>(host.blob) + frag_bytes, tail.data(), tail.size() * sizeof(float),
                  hipMemcpyHostToDevice) != hipSuccess) {
        proposer_release(host);
        return "proposer weight upload failed";
    }
    const float *dtail = reinterpret_cast<const float *>(static_cast<char *>(host.blob) + frag_bytes);
    host.dev.frags = static_cast<const bf16x8 *>(host.blob);
    host.dev.bias = dtail;
    host.dev.w_out = dtail + 256;
    host.dev.b_out = params[kB5];
    host.dev.inv_half = 1.0f / box_half;
    host.dev.box_half = box_half;
    host.dev.switch_distance = switch_distance;
    host.dev.resume_distance = resume_distance;
    host.dev.step_scale = step_scale;
    host.enabled = true;
    return nullptr;
}

void proposer_release(ProposerHost &host) {
    if (host.blob) (void)hipFree(host.blob);
    host.blob = nullptr;
    host.enabled = false;
    host.dev = ProposerDevice{};
}

// One proposal per point; 256-thread blocks, tail lanes evaluate a copy of
// point 0 (the MLP needs the whole wave) and store nothing.
__global__ __launch_bounds__(256) void proposer_eval_kernel(ProposerDevice P, const float4 *__restrict__ pts, size_t n,
                                                            float *__restrict__ out) {
    const size_t i = size_t(blockIdx.x) * 256u + threadIdx.x;
    const float4 q = pts[i < n ? i : 0];
    const float d = proposer_eval(P, mk(q.x, q.y, q.z));
    if (i < n) out[i] = d;
}

hipError_t launch_proposer_eval(const ProposerDevice &P, const float4 *pts, size_t n, float *out, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(proposer_eval_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream, P, pts, n, out);
    return hipGetLastError();
}

// splitmix64 sequence (Steele, Lea, Flood 2014): state += golden gamma, then
// the variant-13 finaliser.
__device__ __forceinline__ uint64_t splitmix_next(uint64_t &state) {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ float splitmix_unit(uint64_t &state) {  // [0, 1), 24 bits
    return float(uint32_t(splitmix_next(state) >> 40)) * (1.0f / 16777216.0f);
}

// Training samples: point i draws candidates uniformly in [-h, h]^3 from the
// sequence seeded with seed ^ (i * 0xD1B54A32D192ED03); the first half of the
// points take the first candidate with a finite distance, the second half the
// first candidate with |d| < band (at most max_attempts candidates; a point
// that finds none keeps its last candidate and counts one failure).  Output
// float4 (x, y, z, d) with d the scene's exact FP32 distance (mapSdf).
__global__ __launch_bounds__(256) void proposer_sample_kernel(const FptRenderConfig *__restrict__ cfg, float half,
                                                              float band, uint32_t n, uint64_t seed,
                                                              uint32_t max_attempts, float4 *__restrict__ out,
                                                              uint32_t *__restrict__ failures) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const bool near = i >= n / 2u;
    uint64_t state = seed ^ (uint64_t(i) * 0xD1B54A32D192ED03ull);
    Work w = {0, 0};
    f3 p = mk(0.f, 0.f, 0.f);
    float d = 0.f;
    bool ok = false;
    for (uint32_t a = 0; a < max_attempts && !ok; ++a) {
        const float ux = splitmix_unit(state), uy = splitmix_unit(state), uz = splitmix_unit(state);
        p = mk((2.0f * ux - 1.0f) * half, (2.0f * uy - 1.0f) * half, (2.0f * uz - 1.0f) * half);
        d = map_scene(p, *cfg, w);
        ok = isfinite(d) && (!near || fabsf(d) < band);
    }
    if (!ok) atomicAdd(failures, 1u);
    out[i] = make_float4(p.x, p.y, p.z, d);
}

hipError_t launch_proposer_samples(const FptRenderConfig *dcfg, float box_half, float band, uint32_t count,
                                   uint64_t seed, uint32_t max_attempts, float4 *out, uint32_t *failures,
                                   hipStream_t stream) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(proposer_sample_kernel, dim3((count + 255u) / 256u), dim3(256), 0, stream, dcfg, box_half, band,
                       count, seed, max_attempts, out, failures);
    return hipGetLastError();
}

}  // namespace fpt
