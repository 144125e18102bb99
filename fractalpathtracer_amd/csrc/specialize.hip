// Topology specialisation of typed SDF programs with the HIP runtime compiler
// (the reference's runtime-source path: generated bodies at the markers
// Shaders.metal:2062-2067 and 3379-3386, generated by
// metal_bridge.mm:592-1300, compiled by metal_bridge.mm:72-95, validated by
// Shaders.metal:5055-5103).
//
// A program's topology is its opcode sequence.  For a topology the library
// generates one translation unit:
//   - topology_distance(): the program's distance as straight-line calls of
//     eval_op<OPCODE> (fpt_device.h), one per instruction, and
//   - topology_surface(): the analytic-gradient surface as surf_op<OPCODE>,
// both reading every numeric parameter (and the CSG flags) from the config
// at run time, so one compiled module serves every program of that topology;
// then it includes the persistent path tracer (persistent_kernel.h) built
// with FPT_TOPOLOGY_SPECIALIZED, whose distance and normal calls go to the
// generated functions.  The interpreter dispatches to the same eval_op /
// surf_op steps, so the specialised kernel evaluates the same float
// operations in the same order.
//
// hipRTC compiles the unit against the library's own device headers, which
// are embedded in the shared library at build time, for the current device's
// architecture.  Code objects are cached in memory per (topology, options,
// device) and on disk under $FPT_HIPRTC_CACHE (default ~/.cache/fpt_hip);
// a plan attaches a module only after a validation kernel has compared the
// specialised distance and gradient with the interpreter's at 2^20 points:
// |d_s - d_i| <= 2e-5 (1 + |d_i|), max gradient component error <= 1e-4.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "fpt_internal.h"
#include "learned.h"
#include "specialize.h"
#include "embedded_headers.h"  // generated at build time: kEmbeddedHeaderNames / kEmbeddedHeaderText

namespace fpt {

namespace {

const char *opcode_supported(uint32_t op) {
    switch (op) {
        case FPT_SDF_OP_ABS: case FPT_SDF_OP_TRANSLATE: case FPT_SDF_OP_SCALE: case FPT_SDF_OP_ROTATE_X:
        case FPT_SDF_OP_ROTATE_Y: case FPT_SDF_OP_ROTATE_Z: case FPT_SDF_OP_REPEAT: case FPT_SDF_OP_SORT_DESC:
        case FPT_SDF_OP_BOX_FOLD: case FPT_SDF_OP_SPHERE: case FPT_SDF_OP_BOX: case FPT_SDF_OP_PLANE:
        case FPT_SDF_OP_MANDELBULB: case FPT_SDF_OP_MANDELBOX: case FPT_SDF_OP_MENGER: case FPT_SDF_OP_ORBIT_ADD:
            return "step";
        case FPT_SDF_OP_UNION: case FPT_SDF_OP_INTERSECTION: case FPT_SDF_OP_SUBTRACT:
            return "marker";  // no effect on the distance (the interpreter skips them too)
        default:
            return nullptr;
    }
}

uint64_t fnv1a(const std::string &s, uint64_t h = 1469598103934665603ull) {
    for (unsigned char c : s) {
        h ^= c;
        h *= 1099511628211ull;
    }
    return h;
}

struct Cache {
    std::mutex mu;
    std::map<std::string, std::shared_ptr<SpecializedModule>> modules;  // key: device + source hash
};
Cache &cache() {
    static Cache *c = new Cache();  // lives until the process ends (modules stay loaded)
    return *c;
}

std::string cache_dir() {
    if (const char *d = std::getenv("FPT_HIPRTC_CACHE")) return d;
    if (const char *x = std::getenv("XDG_CACHE_HOME")) return std::string(x) + "/fpt_hip";
    if (const char *h = std::getenv("HOME")) return std::string(h) + "/.cache/fpt_hip";
    return "";
}

void make_dirs(const std::string &path) {
    std::string cur;
    for (size_t i = 0; i < path.size(); ++i) {
        cur += path[i];
        if (path[i] == '/' && cur.size() > 1) mkdir(cur.c_str(), 0755);
    }
    mkdir(path.c_str(), 0755);
}

bool read_file(const std::string &path, std::vector<char> &out) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    out.resize(n > 0 ? size_t(n) : 0);
    bool ok = n > 0 && std::fread(out.data(), 1, out.size(), f) == out.size();
    std::fclose(f);
    return ok;
}

void write_file_atomic(const std::string &path, const std::vector<char> &data) {
    const std::string tmp = path + ".tmp" + std::to_string(uint64_t(std::chrono::steady_clock::now()
                                                                        .time_since_epoch().count()));
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) return;
    const bool ok = std::fwrite(data.data(), 1, data.size(), f) == data.size();
    std::fclose(f);
    if (ok) std::rename(tmp.c_str(), path.c_str());
    else std::remove(tmp.c_str());
}

const char *const kOptions[] = {"-O3", "-std=c++17", "-ffp-contract=on", "-Xclang", "-target-feature",
                                "-Xclang", "-packed-fp32-ops"};

}  // namespace

std::string topology_source(const FptRenderConfig &cfg, std::string &failure) {
    const uint32_t count = cfg.sdf_program_count < uint32_t(FPT_SDF_PROGRAM_MAX_OPS) ? cfg.sdf_program_count
                                                                                     : uint32_t(FPT_SDF_PROGRAM_MAX_OPS);
    if (cfg.sdf_id != FPT_SDF_PROGRAM || count == 0u) {
        failure = "topology specialization requires a typed geometry program";
        return "";
    }
    std::string dist, surf;
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t op = cfg.sdf_program[i].opcode;
        const char *kind = opcode_supported(op);
        if (!kind) {
            failure = "typed program contains unsupported opcode " + std::to_string(op);
            return "";
        }
        if (!std::strcmp(kind, "marker")) continue;
        const std::string k = std::to_string(op) + "u", idx = std::to_string(i);
        dist += "    eval_op<" + k + ", false>(ops[" + idx + "], s, oc);\n";
        if (op != FPT_SDF_OP_ORBIT_ADD) surf += "    surf_op<" + k + ">(ops[" + idx + "], s);\n";
    }
    std::string src;
    src += "// Generated: topology-specialised typed program (" + std::to_string(count) + " instructions).\n";
    src += "#define FPT_TOPOLOGY_SPECIALIZED 1\n#include \"persistent_kernel.h\"\n\nnamespace fpt {\n\n";
    src += "__device__ float topology_distance(f3 p, const FptRenderConfig &cfg, int *oc) {\n"
           "    EvalState s = {p, 1.0f, kInf, 0.0f, false};\n"
           "    const FptSdfInstruction *ops = cfg.sdf_program;\n";
    src += dist;
    src += "    return s.dist;\n}\n\n";
    src += "__device__ float topology_surface(f3 p, const FptRenderConfig &cfg, f3 &grad) {\n"
           "    SurfState s = surf_begin(p);\n"
           "    const FptSdfInstruction *ops = cfg.sdf_program;\n";
    src += surf;
    src += "    grad = s.grad;\n    return s.dist;\n}\n\n";
    // Validation kernel: interpreter vs specialised at hashed points of the
    // voxel bounds (plus the origin, both corners and the centre).
    src += R"SRC(__device__ __forceinline__ uint32_t validation_hash(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}

extern "C" __global__ __launch_bounds__(256) void fpt_topology_validate(const FptRenderConfig *__restrict__ cfgp,
                                                                     uint32_t n, unsigned int *counts) {
    const FptRenderConfig &cfg = *cfgp;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const f3 lo = ld3(cfg.voxel_bounds_min), hi = ld3(cfg.voxel_bounds_max);
    const float k = 1.0f / 4294967295.0f;
    f3 u = mk(float(validation_hash(i * 3u + 0x9e3779b9u)) * k, float(validation_hash(i * 3u + 0x243f6a88u)) * k,
              float(validation_hash(i * 3u + 0xb7e15162u)) * k);
    f3 p = lo + mk((hi.x - lo.x) * u.x, (hi.y - lo.y) * u.y, (hi.z - lo.z) * u.z);
    if (i == 0u) p = mk(0.f, 0.f, 0.f);
    if (i == 1u) p = lo;
    if (i == 2u) p = hi;
    if (i == 3u) p = (lo + hi) * 0.5f;
    const float di = program_eval<false>(cfg.sdf_program, prog_count(cfg), p, nullptr, nullptr);
    const float ds = topology_distance(p, cfg, nullptr);
    const float de = fabsf(di - ds);
    if (!isfinite(di) || !isfinite(ds) || de > 2.0e-5f * (1.0f + fabsf(di))) atomicAdd(&counts[0], 1u);
    if (isfinite(de)) atomicMax(&counts[2], __float_as_uint(de));
    f3 gi, gs;
    program_surface(cfg.sdf_program, prog_count(cfg), p, gi);
    topology_surface(p, cfg, gs);
    const float ge = fmaxf(fabsf(gi.x - gs.x), fmaxf(fabsf(gi.y - gs.y), fabsf(gi.z - gs.z)));
    if (!finite3(gi) || !finite3(gs) || ge > 1.0e-4f) atomicAdd(&counts[1], 1u);
    if (isfinite(ge)) atomicMax(&counts[3], __float_as_uint(ge));
}

// Specialised distance and analytic gradient at caller points (the test hook
// that compares this code with the CPU oracle): out[i] = {d, grad}.
extern "C" __global__ __launch_bounds__(256) void fpt_topology_sample(const FptRenderConfig *__restrict__ cfgp,
                                                                   const float4 *pts, uint32_t n, float4 *out) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const float4 q = pts[i];
    const f3 p = mk(q.x, q.y, q.z);
    f3 g;
    topology_surface(p, *cfgp, g);
    out[i] = make_float4(topology_distance(p, *cfgp, nullptr), g.x, g.y, g.z);
}

}  // namespace fpt
)SRC";
    return src;
}

int specialize(const FptRenderConfig &cfg, std::shared_ptr<SpecializedModule> &out, SpecializeReport &rep,
               std::string &failure) {
    rep = SpecializeReport();
    const std::string src = topology_source(cfg, failure);
    if (src.empty()) return 1;
    rep.source_bytes = src.size();
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        failure = "no HIP device";
        return 1;
    }
    std::string arch = prop.gcnArchName;  // e.g. "gfx950:sramecc+:xnack-"
    // The key covers everything the code object depends on: the source, the
    // options, the embedded headers and the compiler and driver versions (a
    // cached object from another ROCm release is never loaded).
    int rtc_major = 0, rtc_minor = 0, driver = 0, runtime = 0;
    (void)hiprtcVersion(&rtc_major, &rtc_minor);
    (void)hipDriverGetVersion(&driver);
    (void)hipRuntimeGetVersion(&runtime);
    std::string opts_key = "--offload-arch=" + arch;
    for (const char *o : kOptions) opts_key += std::string(" ") + o;
    opts_key += " hiprtc=" + std::to_string(rtc_major) + "." + std::to_string(rtc_minor) +
                " driver=" + std::to_string(driver) + " runtime=" + std::to_string(runtime);
    uint64_t h = fnv1a(src);
    h = fnv1a(opts_key, h);
    for (size_t i = 0; i < kEmbeddedHeaderCount; ++i) h = fnv1a(kEmbeddedHeaderText[i], h);
    char hex[17];
    std::snprintf(hex, sizeof(hex), "%016llx", (unsigned long long)h);
    const std::string key = std::to_string(dev) + ":" + hex;
    {
        std::lock_guard<std::mutex> lock(cache().mu);
        auto it = cache().modules.find(key);
        if (it != cache().modules.end()) {
            out = it->second;
            rep.cache_status = 2u;
            rep.code_bytes = out->code_bytes;
            return 0;
        }
    }
    auto compile = [&](std::vector<char> &code) -> bool {
        hiprtcProgram prog;
        if (hiprtcCreateProgram(&prog, src.c_str(), "fpt_topology.hip", int(kEmbeddedHeaderCount), kEmbeddedHeaderText,
                                kEmbeddedHeaderNames) != HIPRTC_SUCCESS) {
            failure = "hiprtcCreateProgram failed";
            return false;
        }
        const std::string arch_opt = "--offload-arch=" + arch.substr(0, arch.find(':'));
        std::vector<const char *> opts = {arch_opt.c_str()};
        for (const char *o : kOptions) opts.push_back(o);
        const char *kname = "&fpt::sdf_persistent_kernel<false, false, false, false, false, false>";
        hiprtcAddNameExpression(prog, kname);
        const hiprtcResult r = hiprtcCompileProgram(prog, int(opts.size()), opts.data());
        if (r != HIPRTC_SUCCESS) {
            size_t n = 0;
            hiprtcGetProgramLogSize(prog, &n);
            std::string log(n, '\0');
            if (n) hiprtcGetProgramLog(prog, &log[0]);
            failure = std::string("hipRTC compilation failed: ") + hiprtcGetErrorString(r) + "\n" + log.substr(0, 2000);
            hiprtcDestroyProgram(&prog);
            return false;
        }
        size_t n = 0;
        hiprtcGetCodeSize(prog, &n);
        code.resize(n);
        hiprtcGetCode(prog, code.data());
        hiprtcDestroyProgram(&prog);
        return true;
    };
    // Loads the code object and looks up both kernels (the kernel's mangled
    // name does not depend on the topology).
    auto load = [&](const std::vector<char> &code, SpecializedModule &m) -> bool {
        if (hipModuleLoadData(&m.module, code.data()) != hipSuccess) {
            m.module = nullptr;
            failure = "hipModuleLoadData failed for the specialised code object";
            return false;
        }
        if (hipModuleGetFunction(&m.kernel, m.module,
                                 "_ZN3fpt21sdf_persistent_kernelILb0ELb0ELb0ELb0ELb0ELb0EEEvPK15FptRenderConfigPKNS_"
                                 "10BatchSceneEjP15HIP_vector_typeIfLj4EEjjPjjiiPNS_13DeviceProfileENS_14ProposerDeviceEPKjNS_11SampleSplitE") !=
                hipSuccess ||
            hipModuleGetFunction(&m.validate, m.module, "fpt_topology_validate") != hipSuccess ||
            hipModuleGetFunction(&m.sample, m.module, "fpt_topology_sample") != hipSuccess) {
            (void)hipModuleUnload(m.module);
            m.module = nullptr;
            failure = "specialised kernels are missing from the code object";
            return false;
        }
        return true;
    };
    std::vector<char> code;
    const std::string dir = cache_dir();
    const std::string path = dir.empty() ? std::string() : dir + "/topology-" + hex + ".co";
    const bool from_disk = !path.empty() && read_file(path, code);
    auto t0 = std::chrono::steady_clock::now();
    rep.cache_status = from_disk ? 2u : 1u;
    if (!from_disk) {
        if (!compile(code)) return 1;
        if (!path.empty()) {
            make_dirs(dir);
            write_file_atomic(path, code);
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    auto mod = std::make_shared<SpecializedModule>();
    if (!load(code, *mod)) {
        if (!from_disk) return 1;
        // A truncated or foreign file in the disk cache: drop it and compile
        // once.
        (void)std::remove(path.c_str());
        t0 = std::chrono::steady_clock::now();
        if (!compile(code)) return 1;
        make_dirs(dir);
        write_file_atomic(path, code);
        rep.cache_status = 1u;
        t1 = std::chrono::steady_clock::now();
        if (!load(code, *mod)) return 1;
    }
    rep.compile_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    rep.link_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
    mod->code_bytes = code.size();
    rep.code_bytes = code.size();
    {
        std::lock_guard<std::mutex> lock(cache().mu);
        cache().modules[key] = mod;
    }
    out = mod;
    return 0;
}

hipError_t validate_specialized(const SpecializedModule &m, const FptRenderConfig *dcfg, uint32_t points,
                                SpecializeReport &rep, hipStream_t stream) {
    unsigned int *d = nullptr;
    hipError_t e = hipMalloc(&d, 4 * sizeof(unsigned int));
    if (e != hipSuccess) return e;
    unsigned int host[4] = {0, 0, 0, 0};
    e = hipMemcpyAsync(d, host, sizeof(host), hipMemcpyHostToDevice, stream);
    void *args[] = {const_cast<FptRenderConfig **>(&dcfg), &points, &d};
    auto t0 = std::chrono::steady_clock::now();
    if (e == hipSuccess)
        e = hipModuleLaunchKernel(m.validate, (points + 255u) / 256u, 1, 1, 256, 1, 1, 0, stream, args, nullptr);
    if (e == hipSuccess) e = hipMemcpyAsync(host, d, sizeof(host), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    (void)hipFree(d);
    rep.validate_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    rep.sample_count = points;
    rep.distance_failures = host[0];
    rep.gradient_failures = host[1];
    std::memcpy(&rep.max_distance_error, &host[2], 4);
    std::memcpy(&rep.max_gradient_error, &host[3], 4);
    return e;
}

hipError_t sample_specialized(const SpecializedModule &m, const FptRenderConfig *dcfg, const float4 *pts, uint32_t n,
                              float4 *out, hipStream_t stream) {
    if (n == 0u) return hipSuccess;
    void *args[] = {const_cast<FptRenderConfig **>(&dcfg), const_cast<float4 **>(&pts), &n, &out};
    return hipModuleLaunchKernel(m.sample, (n + 255u) / 256u, 1, 1, 256, 1, 1, 0, stream, args, nullptr);
}

hipError_t launch_specialized(const SpecializedModule &m, const FptRenderConfig *dcfg, float4 *accum, uint32_t start,
                              uint32_t count, uint32_t *counter, uint32_t slots, uint32_t workgroups, int event_batch,
                              int refill, const uint32_t *tile_order, const SampleSplit &split, hipStream_t stream) {
    const BatchScene *scenes = nullptr;
    uint32_t nscenes = 1u;
    DeviceProfile *prof = nullptr;
    ProposerDevice none{};
    void *args[] = {const_cast<FptRenderConfig **>(&dcfg), &scenes, &nscenes, &accum, &start, &count, &counter,
                    &slots, &event_batch, &refill, &prof, &none, const_cast<uint32_t **>(&tile_order),
                    const_cast<SampleSplit *>(&split)};
    return hipModuleLaunchKernel(m.kernel, workgroups, 1, 1, 256, 1, 1, 0, stream, args, nullptr);
}

}  // namespace fpt
