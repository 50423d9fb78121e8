/*
 * jit.cpp - composes, compiles (hipRTC, gfx950) and launches the per-pass
 * kernels, the way the reference composes one OpenCL program from a prelude,
 * the model sources and generated per-pass kernel text and builds it with
 * clBuildProgram (tay/ocl.c:100-177, tay/ocl_wrappers.c:191-211).
 *
 * Program text:
 *   device_abi.h                      argument structures
 *   prelude                           Tay's model dialect: global/constant/kernel
 *                                     qualifiers, packed float2/3/4, TayAgentTag,
 *                                     the agent position accessors
 *                                     (tay/tay.h:4-13, tay/ocl.c:111-126)
 *   model sources, in order           functions made device-callable by
 *                                     clang's force_cuda_host_device pragma
 *   jit_device.h                      see/act templates
 *   one kernel per pass               tay_pass_<i>
 *
 * The compile pins the float semantics the CPU oracle uses: no contraction
 * into fused multiply-add, correctly rounded division and square root,
 * denormals kept, and no packed FP32 instructions.
 */
#include "internal.h"

#include <hip/hiprtc.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <unordered_map>

using namespace tay;

namespace {

const char *const kPrelude = R"TAYPRELUDE(
#pragma clang force_cuda_host_device begin
#define global
#define constant
#define kernel
#define float2 tay_float2
#define float3 tay_float3
#define float4 tay_float4
typedef struct __attribute__((packed)) tay_float2 { union { struct { float x, y; }; float arr[2]; }; } tay_float2;
typedef struct __attribute__((packed)) tay_float3 { union { struct { float x, y, z; }; float arr[3]; }; } tay_float3;
typedef struct __attribute__((packed)) tay_float4 {
    union { struct { float x, y, z, w; }; tay_float3 xyz; float arr[4]; };
} tay_float4;
typedef struct __attribute__((packed)) { tay_float4 min; tay_float4 max; } Box;
typedef struct __attribute__((packed)) TayAgentTag { unsigned part_i; unsigned cell_agent_i; } TayAgentTag;
typedef enum TayBool { TAY_FALSE, TAY_TRUE } TayBool;
#define float4_agent_position(t) (*(float4 *)((TayAgentTag *)(t) + 1))
#define float3_agent_position(t) (*(float3 *)((TayAgentTag *)(t) + 1))
#define float2_agent_position(t) (*(float2 *)((TayAgentTag *)(t) + 1))
#define float4_agent_min(t) (*(float4 *)((TayAgentTag *)(t) + 1))
#define float3_agent_min(t) (*(float3 *)((TayAgentTag *)(t) + 1))
#define float2_agent_min(t) (*(float2 *)((TayAgentTag *)(t) + 1))
#define float4_agent_max(t) (*(float4 *)((char *)(t) + sizeof(TayAgentTag) + sizeof(float4)))
#define float3_agent_max(t) (*(float3 *)((char *)(t) + sizeof(TayAgentTag) + sizeof(float4)))
#define float2_agent_max(t) (*(float2 *)((char *)(t) + sizeof(TayAgentTag) + sizeof(float4)))
__device__ __forceinline__ float tay_sqrt(float x) { return tay_sqrt_rn(x); }
__device__ __forceinline__ double tay_sqrt(double x) { return __builtin_sqrt(x); }
template <class I> __device__ __forceinline__ double tay_sqrt(I x) { return __builtin_sqrt((double)x); }
#define sqrt(x) tay_sqrt(x)
#define sqrtf(x) tay_sqrt_rn((float)(x))
)TAYPRELUDE";

const char *const kPreludeEnd = R"TAYPRELUDE(
#pragma clang force_cuda_host_device end
#undef global
#undef constant
#undef kernel
#undef sqrt
#undef sqrtf
)TAYPRELUDE";

const char *const kOptions[] = {
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-ffp-contract=off",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-gnu-anonymous-struct",
    "-Wno-nested-anon-types",
};

std::mutex g_cache_mutex;
std::unordered_map<std::string, std::vector<char>> g_code_cache;

int narrow_kind(int seer_point, int seen_point) {
    if (seer_point)
        return seen_point ? 0 : 1;   /* TAY_NARROW_PP, TAY_NARROW_PN */
    return seen_point ? 2 : 3;       /* TAY_NARROW_NP, TAY_NARROW_NN */
}

void append(std::string &s, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void append(std::string &s, const char *fmt, ...) {
    char buf[4096];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    s += buf;
}

bool valid_identifier(const char *name) {
    if (!name || !*name)
        return false;
    for (const char *c = name; *c; ++c) {
        bool ok = (*c >= 'a' && *c <= 'z') || (*c >= 'A' && *c <= 'Z') || *c == '_' || (c != name && *c >= '0' && *c <= '9');
        if (!ok)
            return false;
    }
    return true;
}

TayGpuPassShape shape_of(const TayGpuPassDesc &d) {
    TayGpuPassShape sh;
    std::memset(&sh, 0, sizeof(sh));
    sh.type = d.type;
    sh.func_name = d.func_name;
    sh.seer_agent_size = d.seer->agent_size;
    sh.seer_is_point = d.seer->is_point;
    sh.dims = d.seer->dims;
    sh.context_size = d.context_size;
    if (d.type == TAY_GPU_PASS_SEE) {
        sh.seen_agent_size = d.seen->agent_size;
        sh.seen_is_point = d.seen->is_point;
        sh.seen_kind = d.seen->kind;
        sh.skip_self = d.seer == d.seen && !d.self_see;
        sh.see_order = d.see_order;
        sh.split_lanes = d.split_lanes > 1 ? d.split_lanes : 1u;
        sh.split_clear = d.split_clear;
        sh.split_combine = d.split_combine;
    }
    return sh;
}

/* Program text before the generated kernels: ABI structures, the model
 * dialect, the model sources and the device templates. */
std::string compose_head(const char *const *sources, unsigned sources_count) {
    std::string text;
    text += device_abi_text;
    text += device_math_text;
    text += kPrelude;
    for (unsigned i = 0; i < sources_count; ++i) {
        append(text, "\n#line 1 \"tay_source_%u\"\n", i);
        text += sources[i];
        text += "\n";
    }
    text += kPreludeEnd;
    text += "#line 1 \"jit_device.h\"\n";
    text += jit_device_text;
    text += "\n#line 1 \"tay_passes\"\n";
    return text;
}

/* Register budget of the see kernels, as wavefronts per SIMD the compiler
 * must fit.  A plain see kernel is compiled for 8 (at most 64 VGPRs) and, if
 * it then keeps values in scratch, for 7 (72 VGPRs): its compact loop is
 * vector-issue bound and hides its LDS and L2 latency with more wavefronts,
 * but scratch costs HBM traffic and the moves in and out of it.  Measured:
 * sph_forces / sph_density have no scratch at 8 and run 2-5 % faster there
 * than at 7; flock_see keeps 40 bytes per lane at 8 and none at 7, where the
 * 1M-agent see pass moves 125 MB instead of 207 MB at equal time (first
 * steps 4.92 ms at 7, 4.97 at 8; 5.26 at 6).  The persistent batch kernel
 * keeps the compiler's choice: its sparse instances walk per lane, scattered
 * loads that more resident wavefronts only make miss more.  A split see
 * pass stays at 8 whatever it spills: its spills sit around the lane-group
 * loop, not in the filter and compact loops, and the 1M flocking pass split
 * over 4 lanes runs 3.21 ms per step at 8 (128 bytes of scratch), 3.36 at 7
 * (104 bytes), 3.51 at 6, 3.80 at 5.
 * TAY_SEE_WAVES_PER_EU=<n> (1-8) fixes the budget of every see kernel, 0
 * leaves it to the compiler. */
constexpr int kSeeWavesFirst = 8;
constexpr int kSeeWavesSpilled = 7;

/* The budget from TAY_SEE_WAVES_PER_EU, or -1 when it is not set. */
int see_waves_override() {
    if (const char *w = std::getenv("TAY_SEE_WAVES_PER_EU"))
        return std::atoi(w);
    return -1;
}

/* Seers per see tile for a pass over `seers` seers.  One wavefront owns a
 * tile and every seer meets its pairs on one lane, so a small state's
 * 64-seer tiles leave most SIMDs without a wavefront and the pass takes as
 * long as one tile's sweep; smaller tiles spread the state over more
 * wavefronts (each tile's box, and so its sweep, also shrinks) at the price
 * of idle lanes in the compact loop.  64 while that still gives every SIMD
 * two wavefronts, else the largest of 32 and 16 that does, else 16. */
unsigned auto_tile_seers(unsigned seers) {
    static unsigned simds = 0;
    if (!simds) {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
            simds = 4u * (unsigned)cus;
        else
            simds = 1024u;
    }
    for (unsigned s = 64u; s > 16u; s >>= 1)
        if ((seers + s - 1u) / s >= 2u * simds)
            return s;
    return 16u;
}

std::string see_occupancy(int n) {
    if (n >= 1 && n <= 8)
        return " __attribute__((amdgpu_waves_per_eu(" + std::to_string(n) + ")))";
    return std::string();
}

/* waves[i]: register budget of see pass i (see_occupancy) */
std::string compose(const char *const *sources, unsigned sources_count, const TayGpuPassShape *shapes,
                    unsigned shapes_count, int count_pairs, const int *waves) {
    std::string text = compose_head(sources, sources_count);
    for (unsigned i = 0; i < shapes_count; ++i) {
        const TayGpuPassShape &p = shapes[i];
        if (p.type == TAY_GPU_PASS_SEE) {
            append(text, "extern \"C\" __global__ __launch_bounds__(%u)%s", kSeeBlock,
                   see_occupancy(waves[i]).c_str());
            append(text,
                   " void tay_pass_%u(TaySeeArgs args) {\n"
                   "    typedef tay_see_traits<decltype(%s)> T;\n"
                   "    tay_see_tiles<T, %uu, %uu, %d, %s, %s, %d, %s, %uu, %s, %uu>(args,\n"
                   "        [](T::seer *a, T::seen *b, T::context *c) { %s(a, b, c); },\n",
                   i, p.func_name, p.seer_agent_size / 4u, p.seen_agent_size / 4u,
                   narrow_kind(p.seer_is_point, p.seen_is_point), p.skip_self ? "true" : "false",
                   p.seen_kind == TAY_GPU_KIND_Z_GRID ? "true" : "false", p.dims, count_pairs ? "true" : "false",
                   (p.context_size + 3u) / 4u, p.see_order && p.split_lanes <= 1 ? "true" : "false",
                   p.split_lanes > 1 ? p.split_lanes : 1u, p.func_name);
            if (p.split_lanes > 1)
                append(text,
                       "        [](T::seer *a) { %s(a); },\n"
                       "        [](T::seer *a, T::seer *b) { %s(a, b); });\n"
                       "}\n",
                       p.split_clear, p.split_combine);
            else
                append(text,
                       "        [](T::seer *) {}, [](T::seer *, T::seer *) {});\n"
                       "}\n");
        } else {
            append(text,
                   "extern \"C\" __global__ __launch_bounds__(%u) void tay_pass_%u(TayActArgs args) {\n"
                   "    typedef tay_act_traits<decltype(%s)> T;\n"
                   "    tay_act_agents<T, %uu, %uu>(args, [](T::agent *a, T::context *c) { %s(a, c); });\n"
                   "}\n",
                   kActBlock, i, p.func_name, p.seer_agent_size / 4u, (p.context_size + 3u) / 4u, p.func_name);
        }
    }
    return text;
}

}  // namespace

bool tay::compile_program(const std::string &text, std::vector<char> &code, char *log, unsigned log_size) {
    {
        std::lock_guard<std::mutex> lock(g_cache_mutex);
        auto it = g_code_cache.find(text);
        if (it != g_code_cache.end()) {
            code = it->second;
            return true;
        }
    }
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, text.c_str(), "tay_program.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        set_error("hiprtcCreateProgram failed");
        return false;
    }
    const int nopts = (int)(sizeof(kOptions) / sizeof(kOptions[0]));
    hiprtcResult r = hiprtcCompileProgram(prog, nopts, kOptions);
    size_t log_len = 0;
    hiprtcGetProgramLogSize(prog, &log_len);
    std::string build_log(log_len + 1, '\0');
    if (log_len)
        hiprtcGetProgramLog(prog, &build_log[0]);
    if (r != HIPRTC_SUCCESS) {
        if (log && log_size) {
            std::strncpy(log, build_log.c_str(), log_size - 1);
            log[log_size - 1] = '\0';
        }
        set_error("hipRTC compile failed: %s\n%s", hiprtcGetErrorString(r), build_log.c_str());
        hiprtcDestroyProgram(&prog);
        return false;
    }
    size_t size = 0;
    hiprtcGetCodeSize(prog, &size);
    code.resize(size);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    /* TAY_JIT_DUMP=<dir>: keep the program text and code object for inspection */
    if (const char *dir = std::getenv("TAY_JIT_DUMP")) {
        static unsigned serial = 0;
        char path[4096];
        std::snprintf(path, sizeof(path), "%s/tay_program_%u.hip", dir, serial);
        if (FILE *f = std::fopen(path, "w")) {
            std::fwrite(text.data(), 1, text.size(), f);
            std::fclose(f);
        }
        std::snprintf(path, sizeof(path), "%s/tay_program_%u.co", dir, serial++);
        if (FILE *f = std::fopen(path, "wb")) {
            std::fwrite(code.data(), 1, code.size(), f);
            std::fclose(f);
        }
    }
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    g_code_cache[text] = code;
    return true;
}

std::string tay::compose_batch(const char *const *sources, unsigned sources_count, const TayGpuBatchPass *passes,
                               unsigned passes_count, unsigned agent_size, int dims, int count_pairs) {
    std::string text = compose_head(sources, sources_count);
    const int forced = see_waves_override();
    const std::string occupancy = see_occupancy(forced >= 0 ? forced : 0);
    for (unsigned i = 0; i < passes_count; ++i) {
        const TayGpuBatchPass &p = passes[i];
        if (p.type == TAY_GPU_PASS_SEE) {
            append(text, "extern \"C\" __global__ __launch_bounds__(%u)%s", kSeeBlock, occupancy.c_str());
            append(text,
                   " void tay_batch_pass_%u(TayBatchSeeArgs args) {\n"
                   "    typedef tay_see_traits<decltype(%s)> T;\n"
                   "    tay_see_batch<T, %uu, %s, %d, %s, %uu>(args,\n"
                   "        [](T::seer *a, T::seen *b, T::context *c) { %s(a, b, c); });\n"
                   "}\n",
                   i, p.func_name, agent_size / 4u, p.self_see ? "false" : "true", dims,
                   count_pairs ? "true" : "false", (p.context_size + 3u) / 4u, p.func_name);
        } else {
            append(text,
                   "extern \"C\" __global__ __launch_bounds__(%u) void tay_batch_pass_%u(TayBatchActArgs args) {\n"
                   "    typedef tay_act_traits<decltype(%s)> T;\n"
                   "    tay_act_batch<T, %uu, %uu>(args, [](T::agent *a, T::context *c) { %s(a, c); });\n"
                   "}\n",
                   kActBlock, i, p.func_name, agent_size / 4u, (p.context_size + 3u) / 4u, p.func_name);
        }
    }
    return text;
}

bool tay::valid_func_name(const char *name) {
    return valid_identifier(name);
}

namespace {
/* A split see pass: 2, 4 or 8 lanes per seer and its two model functions
 * named (0 or 1 lanes: no split). */
bool valid_split(int type, unsigned lanes, const char *clear, const char *combine, unsigned i) {
    if (type != TAY_GPU_PASS_SEE || lanes <= 1)
        return true;
    if (lanes != 2 && lanes != 4 && lanes != 8) {
        set_error("pass %u: a split see pass has 2, 4 or 8 lanes per seer", i);
        return false;
    }
    if (!valid_identifier(clear) || !valid_identifier(combine)) {
        set_error("pass %u: split clear and combine function names are not C identifiers", i);
        return false;
    }
    return true;
}
}  // namespace

TAY_API int tay_gpu_compile_check(const char *const *sources, unsigned sources_count, const TayGpuPassShape *shapes,
                                  unsigned shapes_count, char *log, unsigned log_size, unsigned long long *code_size) {
    for (unsigned i = 0; i < shapes_count; ++i) {
        const TayGpuPassShape &sh = shapes[i];
        if (!valid_identifier(sh.func_name)) {
            set_error("pass %u: function name is not a C identifier", i);
            return 0;
        }
        if (!valid_split(sh.type, sh.split_lanes, sh.split_clear, sh.split_combine, i))
            return 0;
        if (sh.seer_agent_size % 4 || (sh.type == TAY_GPU_PASS_SEE && sh.seen_agent_size % 4) || sh.dims < 1 ||
            sh.dims > 3) {
            set_error("pass %u: agent sizes must be multiples of 4 and dims 1 to 3", i);
            return 0;
        }
    }
    std::vector<char> code;
    const int forced = see_waves_override();
    std::vector<int> waves(shapes_count, forced >= 0 ? forced : kSeeWavesFirst);
    if (!compile_program(compose(sources, sources_count, shapes, shapes_count, 0, waves.data()), code, log, log_size))
        return 0;
    if (code_size)
        *code_size = code.size();
    return 1;
}

TAY_API void tay_gpu_program_destroy(TayGpuProgram *p) {
    if (!p)
        return;
    (void)tay_gpu_synchronize();
    for (auto &pass : p->passes) {
        device_free(pass.context);
        device_free(pass.pair_count);
        if (pass.context_host)
            (void)hipHostFree(pass.context_host);
        for (hipEvent_t e : pass.ev_free)
            (void)hipEventDestroy(e);
        for (auto &e : pass.ev_pending) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
    }
    if (p->module)
        (void)hipModuleUnload(p->module);
    delete p;
}

TAY_API TayGpuProgram *tay_gpu_program_create(const char *const *sources, unsigned sources_count,
                                              const TayGpuPassDesc *passes, unsigned passes_count, int count_pairs,
                                              char *log, unsigned log_size) {
    for (unsigned i = 0; i < passes_count; ++i) {
        const TayGpuPassDesc &d = passes[i];
        if (!valid_identifier(d.func_name)) {
            set_error("pass %u: function name is not a C identifier", i);
            return nullptr;
        }
        if (!valid_split(d.type, d.split_lanes, d.split_clear, d.split_combine, i))
            return nullptr;
        if (!d.seer || (d.type == TAY_GPU_PASS_SEE && !d.seen)) {
            set_error("pass %u: missing group", i);
            return nullptr;
        }
        if (d.type == TAY_GPU_PASS_SEE && d.seer->dims != d.seen->dims) {
            set_error("groups with different numbers of dimensions not supported");
            return nullptr;
        }
        if (d.type == TAY_GPU_PASS_SEE && d.seer->dims > 3) {
            set_error("see passes support 1 to 3 dimensions");
            return nullptr;
        }
    }
    if (!stream())
        return nullptr;
    std::vector<TayGpuPassShape> shapes(passes_count);
    for (unsigned i = 0; i < passes_count; ++i)
        shapes[i] = shape_of(passes[i]);
    const int forced = see_waves_override();
    std::vector<int> waves(passes_count, forced >= 0 ? forced : kSeeWavesFirst);
    std::vector<char> code;
    if (!compile_program(compose(sources, sources_count, shapes.data(), passes_count, count_pairs, waves.data()), code,
                         log, log_size))
        return nullptr;

    TayGpuProgram *p = new TayGpuProgram();
    p->module = nullptr;
    p->count_pairs = count_pairs;
    p->timing = 0;
    p->width = 64;
    p->lane_walk = TAY_GPU_DEFAULT_LANE_WALK;
    p->tile_seers = 0;
    if (!hip_ok(hipModuleLoadData(&p->module, code.data()), "hipModuleLoadData")) {
        tay_gpu_program_destroy(p);
        return nullptr;
    }
    if (forced < 0) {
        /* unsplit see kernels that keep values in scratch at the first
         * budget are compiled again for the second */
        bool again = false;
        for (unsigned i = 0; i < passes_count; ++i) {
            if (passes[i].type != TAY_GPU_PASS_SEE || passes[i].split_lanes > 1)
                continue;
            char name[64];
            std::snprintf(name, sizeof(name), "tay_pass_%u", i);
            hipFunction_t fn = nullptr;
            int scratch = 0;
            if (!hip_ok(hipModuleGetFunction(&fn, p->module, name), "hipModuleGetFunction") ||
                !hip_ok(hipFuncGetAttribute(&scratch, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, fn), "hipFuncGetAttribute")) {
                tay_gpu_program_destroy(p);
                return nullptr;
            }
            if (scratch > 0) {
                waves[i] = kSeeWavesSpilled;
                again = true;
            }
        }
        if (again) {
            (void)hipModuleUnload(p->module);
            p->module = nullptr;
            if (!compile_program(compose(sources, sources_count, shapes.data(), passes_count, count_pairs,
                                         waves.data()),
                                 code, log, log_size) ||
                !hip_ok(hipModuleLoadData(&p->module, code.data()), "hipModuleLoadData")) {
                tay_gpu_program_destroy(p);
                return nullptr;
            }
        }
    }
    p->passes.resize(passes_count);
    for (unsigned i = 0; i < passes_count; ++i) {
        TayGpuProgram::Pass &pass = p->passes[i];
        std::memset((void *)&pass.desc, 0, sizeof(pass.desc));
        pass.desc = passes[i];
        pass.func_name = passes[i].func_name;
        pass.desc.func_name = pass.func_name.c_str();
        pass.desc.split_lanes = passes[i].split_lanes > 1 ? passes[i].split_lanes : 1u;
        pass.split_clear = pass.desc.split_lanes > 1 ? passes[i].split_clear : "";
        pass.split_combine = pass.desc.split_lanes > 1 ? passes[i].split_combine : "";
        pass.desc.split_clear = pass.split_clear.c_str();
        pass.desc.split_combine = pass.split_combine.c_str();
        pass.kernel = nullptr;
        pass.context = nullptr;
        pass.context_host = nullptr;
        pass.pair_count = nullptr;
        pass.ms = 0.0;
        pass.launches = 0;
        char name[64];
        std::snprintf(name, sizeof(name), "tay_pass_%u", i);
        /* whole words: kernels copy the context word by word */
        const unsigned ctx_bytes = passes[i].context_size ? (passes[i].context_size + 3u) & ~3u : 4u;
        bool ok = hip_ok(hipModuleGetFunction(&pass.kernel, p->module, name), "hipModuleGetFunction") &&
                  device_alloc(&pass.context, ctx_bytes, "pass context") &&
                  hip_ok(hipHostMalloc(&pass.context_host, ctx_bytes, hipHostMallocDefault), "hipHostMalloc") &&
                  hip_ok(hipMemsetAsync(pass.context, 0, ctx_bytes, stream()), "hipMemsetAsync");
        if (ok && passes[i].type == TAY_GPU_PASS_SEE)
            ok = device_alloc((void **)&pass.pair_count, 8ull * TAY_PAIR_SLOTS, "pair counters") &&
                 hip_ok(hipMemsetAsync(pass.pair_count, 0, 8ull * TAY_PAIR_SLOTS, stream()), "hipMemsetAsync");
        if (!ok) {
            tay_gpu_program_destroy(p);
            return nullptr;
        }
    }
    return p;
}

TAY_API int tay_gpu_program_set_context(TayGpuProgram *p, unsigned pass_index, const void *context) {
    if (pass_index >= p->passes.size()) {
        set_error("pass index %u out of range", pass_index);
        return 0;
    }
    TayGpuProgram::Pass &pass = p->passes[pass_index];
    if (!pass.desc.context_size || !context)
        return 1;
    hipStream_t s = stream();
    /* the previous upload from the pinned buffer may still be in flight */
    if (!s || !hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize"))
        return 0;
    std::memcpy(pass.context_host, context, pass.desc.context_size);
    return hip_ok(hipMemcpyAsync(pass.context, pass.context_host, pass.desc.context_size, hipMemcpyHostToDevice, s),
                  "upload pass context");
}

/* Adds the elapsed time of every recorded launch to the pass total and
 * recycles the events. */
static bool resolve_timing(TayGpuProgram::Pass &pass) {
    bool ok = true;
    for (auto &e : pass.ev_pending) {
        float ms = 0.0f;
        ok = ok && hip_ok(hipEventSynchronize(e.second), "hipEventSynchronize") &&
             hip_ok(hipEventElapsedTime(&ms, e.first, e.second), "hipEventElapsedTime");
        pass.ms += ms;
        pass.ev_free.push_back(e.first);
        pass.ev_free.push_back(e.second);
    }
    pass.ev_pending.clear();
    return ok;
}

static bool take_event(TayGpuProgram::Pass &pass, hipEvent_t *e) {
    if (!pass.ev_free.empty()) {
        *e = pass.ev_free.back();
        pass.ev_free.pop_back();
        return true;
    }
    return hip_ok(hipEventCreate(e), "hipEventCreate");
}

TAY_API int tay_gpu_program_run_pass(TayGpuProgram *p, unsigned pass_index) {
    if (pass_index >= p->passes.size()) {
        set_error("pass index %u out of range", pass_index);
        return 0;
    }
    hipStream_t s = stream();
    if (!s)
        return 0;
    TayGpuProgram::Pass &pass = p->passes[pass_index];
    const TayGpuPassDesc &d = pass.desc;
    TayGpuSpace *a = d.seer;
    unsigned grid = 0;
    TaySeeArgs see;
    TayActArgs act;
    void *params[1];
    if (d.type == TAY_GPU_PASS_SEE) {
        TayGpuSpace *b = d.seen;
        if (a->count == 0 || b->count == 0)
            return 1;
        std::memset(&see, 0, sizeof(see));
        see.seer_cols = a->cols[a->cur];
        see.seer_out = a->cols[a->cur];
        see.seer_cap = a->capacity;
        see.seer_count = a->count;
        see.seen_cols = b->cols[b->cur];
        see.seen_cap = b->capacity;
        see.seen_count = b->count;
        see.ranges = b->ranges;
        see.geom = b->geom;
        see.context = pass.context;
        for (int k = 0; k < 4; ++k)
            see.radii[k] = d.radii[k];
        see.pair_count = pass.pair_count;
        /* a split pass holds 64 / lanes seers per wavefront */
        see.tile_seers = d.split_lanes > 1 ? 64u / d.split_lanes
                                           : p->tile_seers ? p->tile_seers : auto_tile_seers(a->count);
        see.tiles = (a->count + see.tile_seers - 1u) / see.tile_seers;
        see.width = p->width < see.tile_seers ? p->width : see.tile_seers;
        see.lane_walk = p->lane_walk;
        grid = (see.tiles + (kSeeBlock / 64u) - 1u) / (kSeeBlock / 64u);
        params[0] = &see;
    } else {
        if (a->count == 0)
            return 1;
        act.cols = a->cols[a->cur];
        act.cap = a->capacity;
        act.count = a->count;
        act.context = pass.context;
        grid = (a->count + kActBlock - 1u) / kActBlock;
        params[0] = &act;
    }
    const unsigned block = d.type == TAY_GPU_PASS_SEE ? kSeeBlock : kActBlock;
    size_t arg_size = d.type == TAY_GPU_PASS_SEE ? sizeof(see) : sizeof(act);
    void *config[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, params[0], HIP_LAUNCH_PARAM_BUFFER_SIZE, &arg_size,
                      HIP_LAUNCH_PARAM_END};
    hipEvent_t ev_beg = nullptr, ev_end = nullptr;
    if (p->timing) {
        if (!take_event(pass, &ev_beg) || !take_event(pass, &ev_end) ||
            !hip_ok(hipEventRecord(ev_beg, s), "hipEventRecord"))
            return 0;
    }
    if (!hip_ok(hipModuleLaunchKernel(pass.kernel, grid, 1, 1, block, 1, 1, 0, s, nullptr, config),
                "hipModuleLaunchKernel"))
        return 0;
    if (p->timing) {
        if (!hip_ok(hipEventRecord(ev_end, s), "hipEventRecord"))
            return 0;
        pass.ev_pending.emplace_back(ev_beg, ev_end);
        ++pass.launches;
    }
    return 1;
}

TAY_API int tay_gpu_program_set_tile_width(TayGpuProgram *p, unsigned width) {
    if (width != 64 && width != 32 && width != 16 && width != 8) {
        set_error("tile width %u: must be 64, 32, 16 or 8", width);
        return 0;
    }
    p->width = width;
    return 1;
}

TAY_API int tay_gpu_program_set_tile_seers(TayGpuProgram *p, unsigned seers) {
    if (seers != 0 && seers != 64 && seers != 32 && seers != 16) {
        set_error("seers per tile %u: must be 0 (by the seer count), 64, 32 or 16", seers);
        return 0;
    }
    p->tile_seers = seers;
    return 1;
}

TAY_API int tay_gpu_program_set_lane_walk(TayGpuProgram *p, unsigned width) {
    if (width > 64) {
        set_error("lane walk width %u: must be at most 64", width);
        return 0;
    }
    p->lane_walk = width;
    return 1;
}

/* The sum of a pass's TAY_PAIR_SLOTS pair counters. */
int tay::read_pair_counters(const unsigned long long *counters, unsigned long long *pairs) {
    unsigned long long host[TAY_PAIR_SLOTS];
    hipStream_t s = stream();
    if (!s || !hip_ok(hipMemcpyAsync(host, counters, sizeof(host), hipMemcpyDeviceToHost, s), "read pair count") ||
        !hip_ok(hipStreamSynchronize(s), "read pair count"))
        return 0;
    *pairs = 0;
    for (unsigned k = 0; k < TAY_PAIR_SLOTS; ++k)
        *pairs += host[k];
    return 1;
}

TAY_API int tay_gpu_program_pair_count(TayGpuProgram *p, unsigned pass_index, unsigned long long *pairs) {
    if (pass_index >= p->passes.size() || !p->passes[pass_index].pair_count) {
        set_error("pass %u is not a see pass", pass_index);
        return 0;
    }
    if (!p->count_pairs) {
        set_error("program was created without pair counting");
        return 0;
    }
    return read_pair_counters(p->passes[pass_index].pair_count, pairs);
}

TAY_API int tay_gpu_program_reset_pair_counts(TayGpuProgram *p) {
    hipStream_t s = stream();
    if (!s)
        return 0;
    for (auto &pass : p->passes)
        if (pass.pair_count && !hip_ok(hipMemsetAsync(pass.pair_count, 0, 8ull * TAY_PAIR_SLOTS, s), "reset"))
            return 0;
    return 1;
}

TAY_API int tay_gpu_program_enable_timing(TayGpuProgram *p, int enable) {
    bool ok = true;
    for (auto &pass : p->passes) {
        ok = resolve_timing(pass) && ok;
        pass.ms = 0.0;
        pass.launches = 0;
    }
    p->timing = enable;
    return ok;
}

TAY_API int tay_gpu_program_pass_time(TayGpuProgram *p, unsigned pass_index, double *ms, unsigned *launches) {
    if (pass_index >= p->passes.size()) {
        set_error("pass index %u out of range", pass_index);
        return 0;
    }
    TayGpuProgram::Pass &pass = p->passes[pass_index];
    if (!resolve_timing(pass))
        return 0;
    *ms = pass.ms;
    *launches = pass.launches;
    return 1;
}
