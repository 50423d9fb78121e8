/*
 * tay_api.cpp - the Tay public API (include/tay.h) for GPU spaces, built on
 * the C-ABI of tay_gpu.h.  The state machine follows the reference's
 * (tay/state.c, tay/state_compile.c): groups own host AoS storage, passes are
 * registered in order, tay_run compiles on demand, pushes agents and pass
 * contexts, runs sort -> passes for every step and fetches all agents back.
 */
/* Plain C++: this file sees Tay's packed float4, not HIP's vector types, so
 * it includes no HIP header and reaches the device only through tay_gpu.h. */
#include "tay.h"
#include "tay_ext.h"

#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#define TAY_API extern "C" __attribute__((visibility("default")))

namespace {

constexpr int kMaxGroups = 8;           /* tay/config.h:8 */
constexpr unsigned kMaxPasses = 32;     /* tay/config.h:9 */
constexpr unsigned kMaxSources = 16;    /* tay/ocl.h:4 */
constexpr unsigned kMaxFuncName = 512;  /* tay/state.h:9 */
constexpr unsigned long long kMB = 1ull << 20;
constexpr unsigned kLibraryDefault = ~0u;  /* a tuning knob left at the GPU library's default */

}  // namespace

struct TayGroup {
    char *storage;
    int storage_pinned;         /* storage from tay_gpu_host_alloc, else calloc */
    unsigned agent_size;
    unsigned capacity;
    int is_point;
    unsigned id;
    int active;

    TaySpaceType type;
    int dims;
    float4 min_part_sizes;
    unsigned long long shared_bytes;
    unsigned count;
    int box_fixed;
    float4 box_min, box_max;
    int order_bits;             /* -1: the library's default */

    TayGpuSpace *gpu;
    int push_agents;
};

struct TayPass {
    int type;   /* TayGpuPassType */
    TayGroup *seer;
    TayGroup *seen;
    char func_name[kMaxFuncName];
    float4 radii;
    int self_see;
    void *context;
    unsigned context_size;
    int has_func;
    unsigned split_lanes;       /* see passes: lanes per seer (tay_set_see_split), 0 or 1: no split */
    char split_clear[kMaxFuncName];
    char split_combine[kMaxFuncName];
};

struct TayState {
    TayGroup groups[kMaxGroups];
    TayPass passes[kMaxPasses];
    unsigned passes_count;
    int is_running;
    TayError error;
    std::string error_message;
    double ms_per_step;
    unsigned next_group_id;
    int recompile;
    int device_enabled;
    char *sources[kMaxSources];
    unsigned sources_count;
    TayGpuProgram *program;
    int count_pairs;
    unsigned tile_width;
    unsigned lane_walk;
    unsigned tile_seers;        /* 0: chosen by the seer count */
    int see_order;              /* TaySeeOrder of every see pass */
    int timing;
    int fetch_mode;
};

namespace {

void set_error(TayState *state, TayError error, const char *message) {
    std::fprintf(stderr, "ERROR: %s\n", message);
    state->error = error;
    state->error_message = message;
}

void set_device_error(TayState *state, const char *what) {
    std::string msg = std::string(what) + ": " + tay_gpu_last_error();
    set_error(state, TAY_ERROR_OCL, msg.c_str());
}

/* The GPU space kind a group runs in, or 0 when the group's space type has
 * no GPU implementation.  After tay_switch_to_ocl the reference runs
 * TAY_CPU_Z_GRID and TAY_CPU_SIMPLE groups on the device
 * (tay/state_compile.c:94-107); this library does the same. */
int gpu_kind(const TayState *state, const TayGroup *g) {
    switch (g->type) {
        case TAY_GPU_Z_GRID: return TAY_GPU_KIND_Z_GRID;
        case TAY_GPU_SIMPLE: return TAY_GPU_KIND_SIMPLE;
        case TAY_CPU_Z_GRID: return state->device_enabled ? TAY_GPU_KIND_Z_GRID : 0;
        case TAY_CPU_SIMPLE: return state->device_enabled ? TAY_GPU_KIND_SIMPLE : 0;
        default: return 0;
    }
}

/* Under TAY_FETCH_ON_DEMAND the device holds the newest agents of every
 * group whose host storage has not been changed since the last push; copy
 * them back before the device spaces go away (simulation end, recompile,
 * switch to host), so nothing computed is lost.  Groups with a pending push
 * keep their host records, which the caller changed. */
int fetch_before_release(TayState *state) {
    if (state->fetch_mode != TAY_FETCH_ON_DEMAND)
        return 1;
    for (TayGroup &g : state->groups) {
        if (!g.active || !g.gpu || g.push_agents)
            continue;
        if (!tay_gpu_space_fetch(g.gpu, g.storage, g.count)) {
            set_device_error(state, "fetch agents");
            return 0;
        }
    }
    return 1;
}

void release_device(TayState *state) {
    tay_gpu_program_destroy(state->program);
    state->program = nullptr;
    for (TayGroup &g : state->groups) {
        tay_gpu_space_destroy(g.gpu);
        g.gpu = nullptr;
    }
}

/* Moves a group's host storage to page-locked memory once the device path
 * is chosen, so pushes and fetches copy at the link's full rate; the whole
 * capacity is copied (records the caller wrote but has not committed yet
 * included).  Storage stays ordinary memory when no page-locked memory can
 * be had. */
void pin_storage(TayGroup &g) {
    if (g.storage_pinned)
        return;
    const size_t bytes = (size_t)(g.capacity ? g.capacity : 1) * (size_t)(g.agent_size ? g.agent_size : 1);
    char *pinned = (char *)tay_gpu_host_alloc(bytes);
    if (!pinned)
        return;
    std::memcpy(pinned, g.storage, bytes);
    std::free(g.storage);
    g.storage = pinned;
    g.storage_pinned = 1;
}

/* tay/state_compile.c:5-175 for the device path: validate groups and
 * passes, create the device spaces and compile the program. */
int state_compile(TayState *state) {
    if (!state->recompile)
        return 1;
    state->recompile = 0;

    for (TayGroup &g : state->groups) {
        if (!g.active)
            continue;
        if (g.type == TAY_GPU_Z_GRID || g.type == TAY_GPU_SIMPLE)
            state->device_enabled = tay_gpu_available() || state->device_enabled;
    }
    for (TayGroup &g : state->groups) {
        if (!g.active)
            continue;
        /* the reference's grids take points only (tay/state_compile.c:25-27);
         * the GPU Z-grid also takes box agents, binned by their min corner */
        const bool grid = g.type == TAY_CPU_GRID || g.type == TAY_CPU_Z_GRID;
        if (!g.is_point && grid) {
            set_error(state, TAY_ERROR_POINT_NONPOINT_MISMATCH, "grid structure can only contain point agents");
            return 0;
        }
        if (g.is_point && g.type == TAY_CPU_AABB_TREE) {
            set_error(state, TAY_ERROR_POINT_NONPOINT_MISMATCH, "aabb tree structure can only contain non-point agents");
            return 0;
        }
        if (!gpu_kind(state, &g)) {
            set_error(state, TAY_ERROR_NOT_IMPLEMENTED,
                      state->device_enabled || g.type == TAY_GPU_Z_GRID || g.type == TAY_GPU_SIMPLE
                          ? "space type has no GPU implementation (GPU Z-grid or GPU simple)"
                          : "CPU spaces are not part of this library; use TAY_GPU_Z_GRID or tay_switch_to_ocl()");
            return 0;
        }
    }

    for (unsigned i = 0; i < state->passes_count; ++i) {
        TayPass &p = state->passes[i];
        if (p.context && !p.context_size) {
            set_error(state, TAY_ERROR_OCL, "ocl pass context set but context size is 0");
            return 0;
        }
        if (!p.context && p.context_size) {
            set_error(state, TAY_ERROR_OCL, "ocl pass context not set but context size is not 0");
            return 0;
        }
        if (p.func_name[0] == '\0') {
            set_error(state, TAY_ERROR_OCL, "ocl pass with no func name set");
            return 0;
        }
        if (!p.seer || !p.seer->active || (p.type == TAY_GPU_PASS_SEE && (!p.seen || !p.seen->active))) {
            set_error(state, TAY_ERROR_GROUP_INDEX_OUT_OF_RANGE, "pass refers to a group that does not exist");
            return 0;
        }
        if (p.type == TAY_GPU_PASS_SEE && p.seer->dims != p.seen->dims) {
            set_error(state, TAY_ERROR_NOT_IMPLEMENTED, "groups with different numbers of dimensions not supported");
            return 0;
        }
    }

    if (!state->device_enabled && !tay_gpu_available()) {
        set_error(state, TAY_ERROR_OCL, "device not found");
        return 0;
    }
    state->device_enabled = 1;

    if (!fetch_before_release(state))
        return 0;
    release_device(state);
    for (TayGroup &g : state->groups) {
        if (!g.active)
            continue;
        pin_storage(g);
        float part[4] = {g.min_part_sizes.x, g.min_part_sizes.y, g.min_part_sizes.z, g.min_part_sizes.w};
        g.gpu = tay_gpu_space_create(gpu_kind(state, &g), g.agent_size, g.capacity, g.is_point, g.dims, part,
                                     g.shared_bytes);
        if (!g.gpu || (g.order_bits >= 0 && !tay_gpu_space_set_order_bits(g.gpu, g.order_bits))) {
            set_device_error(state, "cannot create the group's device space");
            release_device(state);
            return 0;
        }
        if (g.box_fixed) {
            float lo[4] = {g.box_min.x, g.box_min.y, g.box_min.z, g.box_min.w};
            float hi[4] = {g.box_max.x, g.box_max.y, g.box_max.z, g.box_max.w};
            tay_gpu_space_fix_box(g.gpu, lo, hi);
        }
        g.push_agents = 1;
    }

    TayGpuPassDesc desc[kMaxPasses];
    for (unsigned i = 0; i < state->passes_count; ++i) {
        const TayPass &p = state->passes[i];
        std::memset(&desc[i], 0, sizeof(desc[i]));
        desc[i].type = p.type;
        desc[i].func_name = p.func_name;
        desc[i].seer = p.seer->gpu;
        desc[i].seen = p.type == TAY_GPU_PASS_SEE ? p.seen->gpu : nullptr;
        desc[i].radii[0] = p.radii.x;
        desc[i].radii[1] = p.radii.y;
        desc[i].radii[2] = p.radii.z;
        desc[i].radii[3] = p.radii.w;
        desc[i].self_see = p.self_see;
        desc[i].context_size = p.context_size;
        /* the balanced order and split passes apply where the seen group is
         * a GPU Z-grid and the seer group a GPU space of its own (not a CPU
         * Z-grid moved to the device), as the oracle states them */
        const bool tiled = p.type == TAY_GPU_PASS_SEE && p.seen->type == TAY_GPU_Z_GRID &&
                           p.seer->type != TAY_CPU_Z_GRID;
        desc[i].see_order = tiled ? state->see_order : TAY_SEE_ORDER_REFERENCE;
        if (tiled && p.split_lanes > 1) {
            desc[i].split_lanes = p.split_lanes;
            desc[i].split_clear = p.split_clear;
            desc[i].split_combine = p.split_combine;
            desc[i].see_order = TAY_SEE_ORDER_REFERENCE;
        } else {
            desc[i].split_lanes = 1;
        }
    }
    static thread_local char log[1 << 16];
    log[0] = '\0';
    state->program = tay_gpu_program_create(state->sources, state->sources_count, desc, state->passes_count,
                                            state->count_pairs, log, sizeof(log));
    if (!state->program) {
        if (log[0])
            std::fprintf(stderr, "%s\n", log);
        set_device_error(state, "cannot build the program");
        release_device(state);
        return 0;
    }
    if (state->timing)
        tay_gpu_program_enable_timing(state->program, 1);
    tay_gpu_program_set_tile_width(state->program, state->tile_width);
    tay_gpu_program_set_tile_seers(state->program, state->tile_seers);
    if (state->lane_walk != kLibraryDefault)
        tay_gpu_program_set_lane_walk(state->program, state->lane_walk);
    return 1;
}

}  // namespace

TAY_API TayState *tay_create_state() {
    TayState *state = new TayState();
    std::memset(state->groups, 0, sizeof(state->groups));
    std::memset(state->passes, 0, sizeof(state->passes));
    std::memset(state->sources, 0, sizeof(state->sources));
    state->passes_count = 0;
    state->is_running = 0;
    state->error = TAY_ERROR_NONE;
    state->ms_per_step = 0.0;
    state->next_group_id = 0;
    state->recompile = 1;
    state->device_enabled = 0;
    state->sources_count = 0;
    state->program = nullptr;
    state->count_pairs = 0;
    state->tile_width = 64;
    state->tile_seers = 0;
    state->see_order = TAY_SEE_ORDER_REFERENCE;
    state->lane_walk = kLibraryDefault;
    state->timing = 0;
    state->fetch_mode = TAY_FETCH_EVERY_RUN;
    return state;
}

TAY_API void tay_destroy_state(TayState *state) {
    if (!state)
        return;
    release_device(state);
    for (TayGroup &g : state->groups)
        if (g.storage_pinned)
            tay_gpu_host_free(g.storage);
        else
            std::free(g.storage);
    for (unsigned i = 0; i < state->sources_count; ++i)
        std::free(state->sources[i]);
    delete state;
}

TAY_API int tay_switch_to_ocl(TayState *state) {
    if (!state->device_enabled) {
        state->device_enabled = tay_gpu_available();
        state->recompile = 1;
        if (!state->device_enabled)
            set_error(state, TAY_ERROR_OCL, "device not found");
    }
    return state->device_enabled;
}

TAY_API int tay_switch_to_host(TayState *state) {
    if (state->device_enabled) {
        fetch_before_release(state);
        release_device(state);
        state->device_enabled = 0;
        state->recompile = 1;
    }
    return state->device_enabled;
}

TAY_API TayError tay_get_error(TayState *state) {
    return state->error;
}

TAY_API const char *tay_get_error_message(TayState *state) {
    return state->error_message.c_str();
}

/* tay/state.c:73-107 */
TAY_API TayGroup *tay_add_group(TayState *state, unsigned agent_size, unsigned agent_capacity, int is_point) {
    int gi = 0;
    for (; gi < kMaxGroups; ++gi)
        if (!state->groups[gi].active)
            break;
    if (gi == kMaxGroups) {
        set_error(state, TAY_ERROR_GROUP_INDEX_OUT_OF_RANGE, "no free group slot");
        return nullptr;
    }
    TayGroup *g = state->groups + gi;
    std::memset(g, 0, sizeof(*g));
    /* ordinary zeroed memory, as the reference's calloc; it moves to
     * page-locked memory when the state compiles for the device
     * (pin_storage), so creating a group touches no GPU */
    g->storage = (char *)std::calloc(agent_capacity ? agent_capacity : 1, agent_size ? agent_size : 1);
    g->storage_pinned = 0;
    if (!g->storage) {
        char msg[96];
        std::snprintf(msg, sizeof(msg), "cannot allocate host storage for %u agents of %u bytes", agent_capacity,
                      agent_size);
        set_error(state, TAY_ERROR_CAPACITY, msg);
        return nullptr;
    }
    g->agent_size = agent_size;
    g->capacity = agent_capacity;
    g->is_point = is_point;
    g->id = state->next_group_id++;
    g->active = 1;
    g->type = TAY_CPU_SIMPLE;
    g->dims = 3;
    g->min_part_sizes.x = g->min_part_sizes.y = g->min_part_sizes.z = g->min_part_sizes.w = 1.0f;
    g->order_bits = -1;
    state->recompile = 1;
    return g;
}

/* tay/state.c:109-118 */
TAY_API void tay_configure_space(TayState *state, TayGroup *group, TaySpaceType space_type, int space_dims,
                                 float4 min_part_sizes, int shared_size_in_megabytes) {
    group->type = space_type;
    group->dims = space_dims;
    group->min_part_sizes = min_part_sizes;
    group->shared_bytes = shared_size_in_megabytes > 0 ? (unsigned long long)shared_size_in_megabytes * kMB : 0ull;
    group->push_agents = 1;
    state->recompile = 1;
}

/* tay/state.c:120-125; unlike the reference's OpenCL grid (todo:6) the GPU
 * Z-grid honours the fixed box. */
TAY_API void tay_fix_space_box(TayState *state, TayGroup *group, float4 min, float4 max) {
    (void)state;
    group->box_min = min;
    group->box_max = max;
    group->box_fixed = 1;
    if (group->gpu) {
        float lo[4] = {min.x, min.y, min.z, min.w};
        float hi[4] = {max.x, max.y, max.z, max.w};
        tay_gpu_space_fix_box(group->gpu, lo, hi);
    }
}

static TayPass *new_pass(TayState *state) {
    if (state->passes_count >= kMaxPasses) {
        set_error(state, TAY_ERROR_NOT_IMPLEMENTED, "too many passes");
        return nullptr;
    }
    TayPass *p = state->passes + state->passes_count++;
    std::memset(p, 0, sizeof(*p));
    state->recompile = 1;
    return p;
}

/* tay/state.c:135-158 */
TAY_API void tay_add_see(TayState *state, TayGroup *seer_group, TayGroup *seen_group,
                         void (*func)(void *, void *, void *), char *func_name, float4 radii, int self_see,
                         void *context, unsigned context_size) {
    TayPass *p = new_pass(state);
    if (!p)
        return;
    p->type = TAY_GPU_PASS_SEE;
    p->seer = seer_group;
    p->seen = seen_group;
    p->has_func = func != nullptr;
    std::snprintf(p->func_name, kMaxFuncName, "%s", func_name ? func_name : "");
    p->radii = radii;
    p->self_see = self_see;
    p->context = context;
    p->context_size = context_size;
}

TAY_API void tay_add_act(TayState *state, TayGroup *act_group, void (*func)(void *, void *), char *func_name,
                         void *context, unsigned context_size) {
    TayPass *p = new_pass(state);
    if (!p)
        return;
    p->type = TAY_GPU_PASS_ACT;
    p->seer = act_group;
    p->has_func = func != nullptr;
    std::snprintf(p->func_name, kMaxFuncName, "%s", func_name ? func_name : "");
    p->context = context;
    p->context_size = context_size;
}

/* tay/state.c:160-181.  Committing agents also marks the group for a push,
 * so agents added between runs reach the device. */
TAY_API void tay_clear_all_agents(TayState *state, TayGroup *group) {
    (void)state;
    group->count = 0;
    group->push_agents = 1;
}

TAY_API void *tay_get_available_agent(TayState *state, TayGroup *group) {
    if (group->count >= group->capacity) {
        set_error(state, TAY_ERROR_CAPACITY, "group is full");
        return nullptr;
    }
    return group->storage + (size_t)group->agent_size * group->count;
}

TAY_API void tay_commit_available_agent(TayState *state, TayGroup *group) {
    tay_commit_available_agents(state, group, 1);
}

TAY_API void tay_commit_available_agents(TayState *state, TayGroup *group, unsigned count) {
    if ((unsigned long long)group->count + count > group->capacity) {
        set_error(state, TAY_ERROR_CAPACITY, "group is full");
        return;
    }
    group->count += count;
    group->push_agents = 1;
}

TAY_API void *tay_get_agent(TayState *state, TayGroup *group, int index) {
    (void)state;
    return group->storage + (size_t)group->agent_size * (size_t)index;
}

TAY_API void tay_simulation_start(TayState *state) {
    if (state->is_running) {
        set_error(state, TAY_ERROR_STATE_STATUS, "cannot start simulation if it's already running");
        return;
    }
    state->recompile = 1;
    state->is_running = 1;
}

TAY_API void tay_simulation_end(TayState *state) {
    if (!state->is_running) {
        set_error(state, TAY_ERROR_STATE_STATUS, "cannot end simulation if it's not running");
        return;
    }
    state->is_running = 0;
    fetch_before_release(state);
    release_device(state);
    state->recompile = 1;
}

/* tay/state.c:194-323 */
TAY_API int tay_run(TayState *state, int steps) {
    if (!state_compile(state))
        return 0;
    if (!state->is_running) {
        set_error(state, TAY_ERROR_STATE_STATUS, "cannot run simulation steps if simulation is not started");
        return 0;
    }
    const auto beg = std::chrono::steady_clock::now();

    for (TayGroup &g : state->groups) {
        if (!g.active || !g.push_agents)
            continue;
        if (!tay_gpu_space_push(g.gpu, g.storage, g.count)) {
            set_device_error(state, "push agents");
            return 0;
        }
        g.push_agents = 0;
    }
    for (unsigned i = 0; i < state->passes_count; ++i)
        if (!tay_gpu_program_set_context(state->program, i, state->passes[i].context)) {
            set_device_error(state, "push pass context");
            return 0;
        }
    if (state->count_pairs && !tay_gpu_program_reset_pair_counts(state->program)) {
        set_device_error(state, "reset pair counts");
        return 0;
    }
    if (state->timing && !tay_gpu_program_enable_timing(state->program, 1)) {
        set_device_error(state, "reset pass timing");
        return 0;
    }

    for (int step = 0; step < steps; ++step) {
        for (TayGroup &g : state->groups)
            if (g.active && !tay_gpu_space_sort(g.gpu)) {
                set_device_error(state, "sort");
                return 0;
            }
        for (unsigned i = 0; i < state->passes_count; ++i)
            if (!tay_gpu_program_run_pass(state->program, i)) {
                set_device_error(state, "pass");
                return 0;
            }
    }

    if (state->fetch_mode == TAY_FETCH_EVERY_RUN && !tay_fetch_agents(state, nullptr))
        return 0;
    if (!tay_gpu_synchronize()) {
        set_device_error(state, "run");
        return 0;
    }
    const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - beg).count();
    state->ms_per_step = steps > 0 ? t / (double)steps * 1000.0 : 0.0;
    return steps;
}

TAY_API double tay_get_ms_per_step_for_last_run(TayState *state) {
    return state->ms_per_step;
}

TAY_API void tay_log(void *file, char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vfprintf(stdout, fmt, ap);
    va_end(ap);
    std::fflush(stdout);
    if (file) {
        va_start(ap, fmt);
        std::vfprintf((FILE *)file, fmt, ap);
        va_end(ap);
    }
}

/* tay/ocl.c:274-282 */
TAY_API void ocl_add_source(TayState *state, const char *source) {
    if (state->sources_count < kMaxSources && source) {
        char *copy = strdup(source);
        if (!copy) {
            set_error(state, TAY_ERROR_CAPACITY, "cannot copy a model source");
            return;
        }
        state->sources[state->sources_count++] = copy;
        state->recompile = 1;
    }
}

/* The device path needs no host worker threads; the pool calls are kept so
 * that hosts written for the reference link unchanged. */
static unsigned g_threads = 0;

TAY_API void tay_threads_start(unsigned threads_count, unsigned thread_storage_size) {
    (void)thread_storage_size;
    g_threads = threads_count ? (threads_count > 64 ? 64 : threads_count) : 1;
}

TAY_API void tay_threads_stop() {
    g_threads = 0;
}

TAY_API unsigned tay_get_number_of_threads() {
    return g_threads;
}

/* ------------------------------------------------------------ extensions */

TAY_API void tay_set_pair_counting(TayState *state, int enable) {
    if (state->count_pairs != (enable != 0)) {
        state->count_pairs = enable != 0;
        state->recompile = 1;
    }
}

TAY_API int tay_get_pair_count(TayState *state, unsigned pass_index, unsigned long long *pairs) {
    if (!state->program) {
        set_error(state, TAY_ERROR_STATE_STATUS, "no program: run the simulation first");
        return 0;
    }
    if (!tay_gpu_program_pair_count(state->program, pass_index, pairs)) {
        set_device_error(state, "pair count");
        return 0;
    }
    return 1;
}

TAY_API int tay_set_order_bits(TayState *state, TayGroup *group, int bits) {
    if (bits < 0 || bits > 3) {
        set_error(state, TAY_ERROR_NOT_IMPLEMENTED, "order bits must be 0 to 3");
        return 0;
    }
    group->order_bits = bits;
    if (group->gpu && !tay_gpu_space_set_order_bits(group->gpu, bits)) {
        set_device_error(state, "order bits");
        return 0;
    }
    return 1;
}

TAY_API int tay_set_lane_walk(TayState *state, unsigned width) {
    if (width > 64) {
        set_error(state, TAY_ERROR_NOT_IMPLEMENTED, "lane walk width must be at most 64");
        return 0;
    }
    state->lane_walk = width;
    if (state->program)
        tay_gpu_program_set_lane_walk(state->program, width);
    return 1;
}

TAY_API int tay_set_tile_seers(TayState *state, unsigned seers) {
    if (seers != 0 && seers != 64 && seers != 32 && seers != 16) {
        set_error(state, TAY_ERROR_NOT_IMPLEMENTED, "seers per tile must be 0, 64, 32 or 16");
        return 0;
    }
    state->tile_seers = seers;
    if (state->program)
        tay_gpu_program_set_tile_seers(state->program, seers);
    return 1;
}

TAY_API int tay_set_see_order(TayState *state, int order) {
    if (order != TAY_SEE_ORDER_REFERENCE && order != TAY_SEE_ORDER_BALANCED) {
        set_error(state, TAY_ERROR_NOT_IMPLEMENTED, "see order must be TAY_SEE_ORDER_REFERENCE or TAY_SEE_ORDER_BALANCED");
        return 0;
    }
    if (state->see_order != order) {
        state->see_order = order;
        state->recompile = 1;
    }
    return 1;
}

TAY_API int tay_set_see_split(TayState *state, unsigned pass_index, unsigned lanes, void (*clear)(void *),
                              const char *clear_name, void (*combine)(void *, void *), const char *combine_name) {
    (void)clear;
    (void)combine;
    if (pass_index >= state->passes_count || state->passes[pass_index].type != TAY_GPU_PASS_SEE) {
        set_error(state, TAY_ERROR_NOT_IMPLEMENTED, "see split: no see pass at that index");
        return 0;
    }
    if (lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8) {
        set_error(state, TAY_ERROR_NOT_IMPLEMENTED, "see split: lanes per seer must be 1, 2, 4 or 8");
        return 0;
    }
    TayPass &p = state->passes[pass_index];
    if (lanes > 1 && (!clear_name || !combine_name || !clear_name[0] || !combine_name[0])) {
        set_error(state, TAY_ERROR_NOT_IMPLEMENTED, "see split: clear and combine functions must be named");
        return 0;
    }
    p.split_lanes = lanes;
    std::snprintf(p.split_clear, kMaxFuncName, "%s", lanes > 1 ? clear_name : "");
    std::snprintf(p.split_combine, kMaxFuncName, "%s", lanes > 1 ? combine_name : "");
    state->recompile = 1;
    return 1;
}

TAY_API int tay_set_tile_width(TayState *state, unsigned width) {
    if (width != 64 && width != 32 && width != 16 && width != 8) {
        set_error(state, TAY_ERROR_NOT_IMPLEMENTED, "tile width must be 64, 32, 16 or 8");
        return 0;
    }
    state->tile_width = width;
    if (state->program)
        tay_gpu_program_set_tile_width(state->program, width);
    return 1;
}

TAY_API void tay_set_pass_timing(TayState *state, int enable) {
    state->timing = enable != 0;
    if (state->program)
        tay_gpu_program_enable_timing(state->program, state->timing);
}

TAY_API int tay_get_pass_time(TayState *state, unsigned pass_index, double *ms, unsigned *launches) {
    if (!state->program) {
        set_error(state, TAY_ERROR_STATE_STATUS, "no program: run the simulation first");
        return 0;
    }
    if (!tay_gpu_program_pass_time(state->program, pass_index, ms, launches)) {
        set_device_error(state, "pass time");
        return 0;
    }
    return 1;
}

TAY_API TayGpuSpace *tay_get_gpu_space(TayState *state, TayGroup *group) {
    (void)state;
    return group->gpu;
}

TAY_API void tay_set_fetch_mode(TayState *state, int mode) {
    state->fetch_mode = mode == TAY_FETCH_ON_DEMAND ? TAY_FETCH_ON_DEMAND : TAY_FETCH_EVERY_RUN;
}

/* tay/ocl.c:261-272, for one group or all */
TAY_API int tay_fetch_agents(TayState *state, TayGroup *group) {
    for (TayGroup &g : state->groups) {
        if (!g.active || !g.gpu || (group && group != &g))
            continue;
        if (!tay_gpu_space_fetch(g.gpu, g.storage, g.count)) {
            set_device_error(state, "fetch agents");
            return 0;
        }
    }
    return 1;
}

TAY_API unsigned tay_fetch_agent_fields(TayState *state, TayGroup *group, unsigned offset, unsigned size,
                                        void *dst) {
    if (offset % 4u || size % 4u || offset > group->agent_size || size > group->agent_size - offset) {
        set_error(state, TAY_ERROR_NOT_IMPLEMENTED, "fields must be whole words within the agent");
        return 0;
    }
    if (!group->gpu) {
        set_error(state, TAY_ERROR_STATE_STATUS, "group is not on the device: run the simulation first");
        return 0;
    }
    const unsigned n = tay_gpu_space_count(group->gpu);
    if (!tay_gpu_space_fetch_words(group->gpu, offset / 4u, size / 4u, dst, n)) {
        set_device_error(state, "fetch agent fields");
        return 0;
    }
    return n;
}
