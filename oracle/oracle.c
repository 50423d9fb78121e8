/*
 * oracle.c - TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Clean-room C restatement of the reference Tay library's CPU path for the
 * Z-grid hot path.  Each function names the reference file:line it follows.
 * Records are AoS, the sort runs serially on the calling thread, see and act
 * passes run on a POSIX thread pool with the reference's static range split,
 * and every pair costs an indirect call, as in the reference.
 *
 * Built with -ffp-contract=off so that no multiply-add is fused: the float
 * arithmetic is the plain IEEE sequence the GPU kernels reproduce.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <semaphore.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#define MAX_GROUPS 8        /* tay/config.h:8 */
#define MAX_PASSES 32       /* tay/config.h:9 */
#define MAX_THREADS 64      /* tay/config.h:10 */
#define MAX_SOURCES 16      /* tay/ocl.h:4 */
#define MAX_FUNC_NAME 512   /* tay/state.h:9 */
#define MB (1 << 20)

/* bytes per cell the reference structures charge against the shared memory */
#define CPU_GRID_CELL_BYTES 32u   /* GridCell on LP64: next, count, first, int4 (tay/cpu_grid.c:7-12) */
#define CPU_Z_GRID_CELL_BYTES 8u  /* ZGridCell: count, first (tay/cpu_z_grid.c:7-10) */
#define GPU_Z_GRID_CELL_BYTES 8u  /* cell range: begin, end */

typedef void (*SeeFunc)(void *, void *, void *);
typedef void (*ActFunc)(void *, void *);

typedef struct { int arr[4]; } Int4;

typedef struct {
    char *agents;
    unsigned size;
    unsigned beg;
    unsigned end;
} Slice;

typedef struct {
    unsigned long long broad;
    unsigned long long narrow;
} PairCounts;

/* ---------------------------------------------------------------- spaces */

typedef struct GridCell {
    struct GridCell *next;
    unsigned count;
    unsigned first;
    Int4 indices;
} GridCell;

typedef struct {
    unsigned count;
    unsigned first;
} ZCell;

typedef struct {
    unsigned beg;
    unsigned end;
} CellRange;

typedef struct {
    TaySpaceType type;
    int dims;
    float4 min_part_sizes;
    unsigned count;
    Box box;
    int is_box_fixed;
    void *shared;
    size_t shared_size;

    /* geometry of the last sort (grid spaces) */
    float4 origin;
    float4 cell_sizes;
    Int4 cell_counts;
    unsigned long long cells_count;

    /* CpuGrid */
    GridCell *grid_cells;
    GridCell *first_cell;
    unsigned long long grid_max_cells;

    /* CpuZGrid */
    ZCell *z_cells;
    unsigned long long z_max_cells;

    /* GPU Z-grid statement */
    CellRange *ranges;
    unsigned long long gz_max_cells;
    uint64_t *sort_buf;     /* (key << 32 | index) pairs */
    int order_bits;         /* requested sub-cell bits of the storage order */
    int gz_order_bits;      /* the ones used by the last sort */
    float gz_max_extent[4]; /* largest max - min corner per axis (non-point agents) */
} Space;

typedef struct TayGroup {
    char *storage;
    char *sort_storage;
    int agent_size;
    unsigned capacity;
    int is_point;
    Space space;
    unsigned id;
    int active;
} TayGroup;

typedef enum { PASS_SEE, PASS_ACT } PassType;

struct TayPass;
typedef void (*PairingFunc)(struct TayPass *, Slice, Slice, PairCounts *);

typedef struct TayPass {
    PassType type;
    TayGroup *seer_group;   /* act group for act passes */
    TayGroup *seen_group;
    SeeFunc see;
    ActFunc act;
    float4 radii;
    int self_see;
    void *context;
    unsigned context_size;
    char func_name[MAX_FUNC_NAME];
    PairingFunc pairing;
    PairCounts counts;      /* summed over the steps of the last run */
    unsigned seer_limit;    /* see passes: only the first seer_limit seers (0: all) */
    /* GPU Z-grid seen spaces: see order and tile layout (the state's) */
    int see_order;
    unsigned tile_seers, tile_width, lane_walk;
    /* split see pass (tay_set_see_split): lanes per seer, partial clear and merge */
    unsigned split_lanes;
    void (*split_clear)(void *);
    void (*split_combine)(void *, void *);
} TayPass;

typedef struct TayState {
    TayGroup groups[MAX_GROUPS];
    TayPass passes[MAX_PASSES];
    unsigned passes_count;
    int is_running;
    TayError error;
    double ms_per_step;
    double see_seconds;
    unsigned next_group_id;
    int recompile;
    char *sources[MAX_SOURCES];
    unsigned sources_count;
    int see_order;          /* 0: reference sweep order, 1: balanced */
    unsigned tile_seers;    /* 0: by the seer count */
    unsigned tile_width;
    unsigned lane_walk;
} TayState;

/* ------------------------------------------------------------ thread pool */
/* POSIX restatement of the reference's semaphore-driven pool
 * (tay/thread.c:13-135): each worker waits on its begin semaphore, runs its
 * task and posts its end semaphore; the caller releases all and waits. */

typedef struct {
    void (*func)(void *task, int thread_i, PairCounts *counts);
    void *task;
} ThreadJob;

typedef struct {
    pthread_t thread;
    sem_t beg, end;
    int run;
    int index;
    ThreadJob job;
    PairCounts counts;
} Worker;

static struct {
    Worker workers[MAX_THREADS];
    int count;
} pool;

static void *worker_main(void *arg) {
    Worker *w = (Worker *)arg;
    for (;;) {
        sem_wait(&w->beg);
        if (!w->run) {
            sem_post(&w->end);
            return 0;
        }
        w->job.func(w->job.task, w->index, &w->counts);
        sem_post(&w->end);
    }
    return 0;
}

void tay_threads_start(unsigned threads_count, unsigned thread_storage_size) {
    (void)thread_storage_size;
    if (pool.count)
        return;
    /* 0 = all processors, capped at the reference's maximum (tay/thread.c:70-76) */
    if (threads_count == 0)
        threads_count = (unsigned)sysconf(_SC_NPROCESSORS_ONLN);
    if (threads_count > MAX_THREADS)
        threads_count = MAX_THREADS;
    if (threads_count == 0)
        threads_count = 1;
    for (unsigned i = 0; i < threads_count; ++i) {
        Worker *w = pool.workers + i;
        w->run = 1;
        w->index = (int)i;
        sem_init(&w->beg, 0, 0);
        sem_init(&w->end, 0, 0);
        pthread_create(&w->thread, 0, worker_main, w);
    }
    pool.count = (int)threads_count;
}

void tay_threads_stop() {
    for (int i = 0; i < pool.count; ++i) {
        pool.workers[i].run = 0;
        sem_post(&pool.workers[i].beg);
    }
    for (int i = 0; i < pool.count; ++i) {
        sem_wait(&pool.workers[i].end);
        pthread_join(pool.workers[i].thread, 0);
        sem_destroy(&pool.workers[i].beg);
        sem_destroy(&pool.workers[i].end);
    }
    pool.count = 0;
}

unsigned tay_get_number_of_threads() {
    return (unsigned)pool.count;
}

static int runner_count(void) {
    return pool.count ? pool.count : 1;
}

/* Static contiguous split, remainder to the last thread (tay/thread.c:244-250). */
static void thread_range(unsigned count, int thread_i, unsigned *beg, unsigned *end) {
    unsigned n = (unsigned)runner_count();
    unsigned per = count / n;
    *beg = per * (unsigned)thread_i;
    *end = ((unsigned)thread_i < n - 1u) ? *beg + per : count;
}

/* Run func on every thread and sum the per-thread pair counters into out
 * (tay/space.c:265-276 + tay/thread.c:78-89).  Without a started pool the
 * single task runs on the caller. */
static void run_tasks(void (*func)(void *, int, PairCounts *), void *task, PairCounts *out) {
    if (pool.count == 0) {
        PairCounts c = {0, 0};
        func(task, 0, &c);
        if (out) {
            out->broad += c.broad;
            out->narrow += c.narrow;
        }
        return;
    }
    for (int i = 0; i < pool.count; ++i) {
        Worker *w = pool.workers + i;
        w->job.func = func;
        w->job.task = task;
        w->counts.broad = w->counts.narrow = 0;
        sem_post(&w->beg);
    }
    for (int i = 0; i < pool.count; ++i)
        sem_wait(&pool.workers[i].end);
    if (out)
        for (int i = 0; i < pool.count; ++i) {
            out->broad += pool.workers[i].counts.broad;
            out->narrow += pool.workers[i].counts.narrow;
        }
}

/* ------------------------------------------------------------------ state */

static void set_error(TayState *state, TayError error, const char *message) {
    fprintf(stderr, "ERROR: %s\n", message);
    state->error = error;
}

TayState *tay_create_state() {
    TayState *state = (TayState *)calloc(1, sizeof(TayState));
    if (state) {
        state->tile_width = 64;     /* the GPU library's defaults */
        state->lane_walk = 64;
    }
    return state;
}

void tay_destroy_state(TayState *state) {
    if (!state)
        return;
    for (int i = 0; i < MAX_GROUPS; ++i) {
        TayGroup *g = state->groups + i;
        free(g->storage);
        free(g->sort_storage);
        free(g->space.shared);
        free(g->space.sort_buf);
    }
    for (unsigned i = 0; i < state->sources_count; ++i)
        free(state->sources[i]);
    free(state);
}

int tay_switch_to_ocl(TayState *state) {
    (void)state;
    return 0;   /* the oracle is CPU only */
}

int tay_switch_to_host(TayState *state) {
    (void)state;
    return 0;
}

TayError tay_get_error(TayState *state) {
    return state->error;
}

/* tay/state.c:73-107 */
TayGroup *tay_add_group(TayState *state, unsigned agent_size, unsigned agent_capacity, int is_point) {
    int gi = 0;
    for (; gi < MAX_GROUPS; ++gi)
        if (!state->groups[gi].active)
            break;
    if (gi == MAX_GROUPS) {
        set_error(state, TAY_ERROR_GROUP_INDEX_OUT_OF_RANGE, "no free group slot");
        return 0;
    }
    TayGroup *g = state->groups + gi;
    memset(g, 0, sizeof(*g));
    g->agent_size = (int)agent_size;
    g->storage = (char *)calloc(agent_capacity ? agent_capacity : 1, agent_size);
    g->sort_storage = (char *)calloc(agent_capacity ? agent_capacity : 1, agent_size);
    g->capacity = agent_capacity;
    g->is_point = is_point;
    g->id = state->next_group_id++;
    g->active = 1;
    g->space.type = TAY_CPU_SIMPLE;
    g->space.dims = 3;
    g->space.min_part_sizes = (float4){{{1.0f, 1.0f, 1.0f, 1.0f}}};
    g->space.order_bits = 2;   /* the GPU library's default */
    return g;
}

/* tay/state.c:109-118 */
void tay_configure_space(TayState *state, TayGroup *group, TaySpaceType space_type, int space_dims, float4 min_part_sizes, int shared_size_in_megabytes) {
    Space *s = &group->space;
    s->type = space_type;
    s->min_part_sizes = min_part_sizes;
    s->dims = space_dims;
    s->shared_size = (size_t)shared_size_in_megabytes * MB;
    s->shared = realloc(s->shared, s->shared_size ? s->shared_size : 1);
    state->recompile = 1;
}

/* tay/state.c:120-125 */
void tay_fix_space_box(TayState *state, TayGroup *group, float4 min, float4 max) {
    (void)state;
    group->space.box.min = min;
    group->space.box.max = max;
    group->space.is_box_fixed = 1;
}

/* tay/state.c:135-158 */
void tay_add_see(TayState *state, TayGroup *seer_group, TayGroup *seen_group, void (*func)(void *, void *, void *), char *func_name, float4 radii, int self_see, void *context, unsigned context_size) {
    if (state->passes_count >= MAX_PASSES) {
        set_error(state, TAY_ERROR_NOT_IMPLEMENTED, "too many passes");
        return;
    }
    TayPass *p = state->passes + state->passes_count++;
    memset(p, 0, sizeof(*p));
    p->type = PASS_SEE;
    p->context = context;
    p->context_size = context_size;
    p->see = func;
    p->radii = radii;
    p->self_see = self_see;
    p->seer_group = seer_group;
    p->seen_group = seen_group;
    snprintf(p->func_name, MAX_FUNC_NAME, "%s", func_name ? func_name : "");
}

void tay_add_act(TayState *state, TayGroup *act_group, void (*func)(void *, void *), char *func_name, void *context, unsigned context_size) {
    if (state->passes_count >= MAX_PASSES) {
        set_error(state, TAY_ERROR_NOT_IMPLEMENTED, "too many passes");
        return;
    }
    TayPass *p = state->passes + state->passes_count++;
    memset(p, 0, sizeof(*p));
    p->type = PASS_ACT;
    p->context = context;
    p->context_size = context_size;
    p->act = func;
    p->seer_group = act_group;
    snprintf(p->func_name, MAX_FUNC_NAME, "%s", func_name ? func_name : "");
}

/* tay/state.c:160-181 */
void tay_clear_all_agents(TayState *state, TayGroup *group) {
    (void)state;
    group->space.count = 0;
}

void *tay_get_available_agent(TayState *state, TayGroup *group) {
    if (group->space.count >= group->capacity) {
        set_error(state, TAY_ERROR_CAPACITY, "group is full");
        return 0;
    }
    return group->storage + (size_t)group->agent_size * group->space.count;
}

void tay_commit_available_agent(TayState *state, TayGroup *group) {
    if (group->space.count >= group->capacity) {
        set_error(state, TAY_ERROR_CAPACITY, "group is full");
        return;
    }
    ++group->space.count;
}

void *tay_get_agent(TayState *state, TayGroup *group, int index) {
    (void)state;
    return group->storage + (size_t)group->agent_size * (size_t)index;
}

void tay_simulation_start(TayState *state) {
    if (state->is_running) {
        set_error(state, TAY_ERROR_STATE_STATUS, "cannot start simulation if it's already running");
        return;
    }
    state->recompile = 1;
    state->is_running = 1;
}

void tay_simulation_end(TayState *state) {
    if (!state->is_running) {
        set_error(state, TAY_ERROR_STATE_STATUS, "cannot end simulation if it's not running");
        return;
    }
    state->is_running = 0;
}

double tay_get_ms_per_step_for_last_run(TayState *state) {
    return state->ms_per_step;
}

void tay_log(void *file, char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vfprintf(stdout, fmt, ap);
    va_end(ap);
    fflush(stdout);
    if (file) {
        va_start(ap, fmt);
        vfprintf((FILE *)file, fmt, ap);
        va_end(ap);
    }
}

void ocl_add_source(TayState *state, const char *source) {
    if (state->sources_count < MAX_SOURCES)
        state->sources[state->sources_count++] = strdup(source);
}

/* ------------------------------------------------------------------- box */

/* tay/space.c:9-48 */
static void update_box(TayGroup *g) {
    Space *s = &g->space;
    if (s->is_box_fixed)
        return;
    for (int i = 0; i < s->dims; ++i) {
        s->box.min.arr[i] = FLT_MAX;
        s->box.max.arr[i] = -FLT_MAX;
    }
    for (unsigned a = 0; a < s->count; ++a) {
        char *agent = g->storage + (size_t)g->agent_size * a;
        float4 lo = float4_agent_min(agent);
        float4 hi = g->is_point ? lo : float4_agent_max(agent);
        for (int i = 0; i < s->dims; ++i) {
            if (lo.arr[i] < s->box.min.arr[i])
                s->box.min.arr[i] = lo.arr[i];
            if (hi.arr[i] > s->box.max.arr[i])
                s->box.max.arr[i] = hi.arr[i];
        }
    }
}

/* -------------------------------------------------------------- pairing */
/* The six pairing loops of tay/space.c:106-238.  Narrow-phase box tests are
 * inclusive per axis: a point pair is skipped when d < -r or d > r for
 * d = seer - seen on any axis (tay/space.c:62-68). */

static inline int narrow_point_point(const float *a, const float *b, const float *r, int dims) {
    for (int i = 0; i < dims; ++i) {
        float d = a[i] - b[i];
        if (d < -r[i] || d > r[i])
            return 0;
    }
    return 1;
}

static inline int narrow_nonpoint_point(const float *amin, const float *amax, const float *b, const float *r, int dims) {
    for (int i = 0; i < dims; ++i) {
        if (amin[i] - b[i] > r[i])
            return 0;
        if (b[i] - amax[i] > r[i])
            return 0;
    }
    return 1;
}

static inline int narrow_point_nonpoint(const float *a, const float *bmin, const float *bmax, const float *r, int dims) {
    for (int i = 0; i < dims; ++i) {
        if (a[i] - bmax[i] > r[i])
            return 0;
        if (bmin[i] - a[i] > r[i])
            return 0;
    }
    return 1;
}

static inline int narrow_nonpoint_nonpoint(const float *amin, const float *amax, const float *bmin, const float *bmax, const float *r, int dims) {
    for (int i = 0; i < dims; ++i) {
        if (amin[i] - bmax[i] > r[i])
            return 0;
        if (bmin[i] - amax[i] > r[i])
            return 0;
    }
    return 1;
}

#define AGENT_AT(slice, i) ((slice).agents + (size_t)(slice).size * (i))
#define POS(agent) ((const float *)((const char *)(agent) + sizeof(TayAgentTag)))
#define MAXC(agent) ((const float *)((const char *)(agent) + sizeof(TayAgentTag) + sizeof(float4)))

static void pair_point_point(TayPass *p, Slice seer, Slice seen, PairCounts *c) {
    int dims = p->seer_group->space.dims;
    for (unsigned i = seer.beg; i < seer.end; ++i) {
        char *a = AGENT_AT(seer, i);
        for (unsigned j = seen.beg; j < seen.end; ++j) {
            if (i == j)
                continue;
            char *b = AGENT_AT(seen, j);
            ++c->broad;
            if (!narrow_point_point(POS(a), POS(b), p->radii.arr, dims))
                continue;
            ++c->narrow;
            p->see(a, b, p->context);
        }
    }
}

static void pair_point_point_self_see(TayPass *p, Slice seer, Slice seen, PairCounts *c) {
    int dims = p->seer_group->space.dims;
    for (unsigned i = seer.beg; i < seer.end; ++i) {
        char *a = AGENT_AT(seer, i);
        for (unsigned j = seen.beg; j < seen.end; ++j) {
            char *b = AGENT_AT(seen, j);
            ++c->broad;
            if (!narrow_point_point(POS(a), POS(b), p->radii.arr, dims))
                continue;
            ++c->narrow;
            p->see(a, b, p->context);
        }
    }
}

static void pair_nonpoint_point(TayPass *p, Slice seer, Slice seen, PairCounts *c) {
    int dims = p->seer_group->space.dims;
    for (unsigned i = seer.beg; i < seer.end; ++i) {
        char *a = AGENT_AT(seer, i);
        for (unsigned j = seen.beg; j < seen.end; ++j) {
            char *b = AGENT_AT(seen, j);
            ++c->broad;
            if (!narrow_nonpoint_point(POS(a), MAXC(a), POS(b), p->radii.arr, dims))
                continue;
            ++c->narrow;
            p->see(a, b, p->context);
        }
    }
}

static void pair_point_nonpoint(TayPass *p, Slice seer, Slice seen, PairCounts *c) {
    int dims = p->seer_group->space.dims;
    for (unsigned i = seer.beg; i < seer.end; ++i) {
        char *a = AGENT_AT(seer, i);
        for (unsigned j = seen.beg; j < seen.end; ++j) {
            char *b = AGENT_AT(seen, j);
            ++c->broad;
            if (!narrow_point_nonpoint(POS(a), POS(b), MAXC(b), p->radii.arr, dims))
                continue;
            ++c->narrow;
            p->see(a, b, p->context);
        }
    }
}

static void pair_nonpoint_nonpoint(TayPass *p, Slice seer, Slice seen, PairCounts *c) {
    int dims = p->seer_group->space.dims;
    for (unsigned i = seer.beg; i < seer.end; ++i) {
        char *a = AGENT_AT(seer, i);
        for (unsigned j = seen.beg; j < seen.end; ++j) {
            if (i == j)
                continue;
            char *b = AGENT_AT(seen, j);
            ++c->broad;
            if (!narrow_nonpoint_nonpoint(POS(a), MAXC(a), POS(b), MAXC(b), p->radii.arr, dims))
                continue;
            ++c->narrow;
            p->see(a, b, p->context);
        }
    }
}

static void pair_nonpoint_nonpoint_self_see(TayPass *p, Slice seer, Slice seen, PairCounts *c) {
    int dims = p->seer_group->space.dims;
    for (unsigned i = seer.beg; i < seer.end; ++i) {
        char *a = AGENT_AT(seer, i);
        for (unsigned j = seen.beg; j < seen.end; ++j) {
            char *b = AGENT_AT(seen, j);
            ++c->broad;
            if (!narrow_nonpoint_nonpoint(POS(a), MAXC(a), POS(b), MAXC(b), p->radii.arr, dims))
                continue;
            ++c->narrow;
            p->see(a, b, p->context);
        }
    }
}

/* ------------------------------------------------------------ cell index */

/* (int)floorf((p - origin) / size) per axis (tay/cpu_grid.c:18-41) */
static inline Int4 cell_indices(const float *p, float4 o, float4 cs, int dims) {
    Int4 idx = {{0, 0, 0, 0}};
    for (int i = 0; i < dims; ++i)
        idx.arr[i] = (int)floorf((p[i] - o.arr[i]) / cs.arr[i]);
    return idx;
}

/* row-major, x fastest (tay/cpu_grid.c:43-64) */
static inline unsigned long long row_major(Int4 idx, Int4 counts, int dims) {
    unsigned long long id = 0, stride = 1;
    for (int i = 0; i < dims; ++i) {
        id += (unsigned long long)idx.arr[i] * stride;
        stride *= (unsigned long long)counts.arr[i];
    }
    return id;
}

/* Morton encode (tay/cpu_z_grid.c:37-83): spread each index with the
 * standard part1by1 / part1by2 bit masks, then interleave. */
static inline unsigned spread2(unsigned v) {
    v &= 0x0000FFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

static inline unsigned spread3(unsigned v) {
    v &= 0x000003FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

static inline unsigned compact2(unsigned v) {
    v &= 0x55555555u;
    v = (v | (v >> 1)) & 0x33333333u;
    v = (v | (v >> 2)) & 0x0F0F0F0Fu;
    v = (v | (v >> 4)) & 0x00FF00FFu;
    v = (v | (v >> 8)) & 0x0000FFFFu;
    return v;
}

static inline unsigned compact3(unsigned v) {
    v &= 0x09249249u;
    v = (v | (v >> 2)) & 0x030C30C3u;
    v = (v | (v >> 4)) & 0x0300F00Fu;
    v = (v | (v >> 8)) & 0xFF0000FFu;
    v = (v | (v >> 16)) & 0x000003FFu;
    return v;
}

unsigned oracle_morton_key(int x, int y, int z, int dims) {
    if (dims == 1)
        return (unsigned)x;
    if (dims == 2)
        return spread2((unsigned)x) | (spread2((unsigned)y) << 1);
    if (dims == 3)
        return spread3((unsigned)x) | (spread3((unsigned)y) << 1) | (spread3((unsigned)z) << 2);
    return 0;
}

void oracle_morton_decode(unsigned key, int dims, int *x, int *y, int *z) {
    *x = *y = *z = 0;
    if (dims == 1) {
        *x = (int)key;
    } else if (dims == 2) {
        *x = (int)compact2(key);
        *y = (int)compact2(key >> 1);
    } else if (dims == 3) {
        *x = (int)compact3(key);
        *y = (int)compact3(key >> 1);
        *z = (int)compact3(key >> 2);
    }
}

static inline unsigned morton(Int4 idx, int dims) {
    return oracle_morton_key(idx.arr[0], idx.arr[1], idx.arr[2], dims);
}

/* --------------------------------------------------------------- CpuGrid */

/* tay/cpu_grid.c:108-114 */
static void cpu_grid_start(Space *s) {
    s->grid_cells = (GridCell *)s->shared;
    s->grid_max_cells = s->shared_size / CPU_GRID_CELL_BYTES;
    for (unsigned long long i = 0; i < s->grid_max_cells; ++i)
        s->grid_cells[i].count = 0;
}

/* tay/cpu_grid.c:116-193: geometry, binning into a prepend list of occupied
 * cells, offsets in list order, scatter, swap. */
static void cpu_grid_sort(TayGroup *g) {
    Space *s = &g->space;
    update_box(g);
    s->first_cell = 0;
    s->origin = s->box.min;

    unsigned long long cells = 1;
    for (int i = 0; i < s->dims; ++i) {
        float part = s->min_part_sizes.arr[i];
        float size = s->box.max.arr[i] - s->box.min.arr[i] + part * 0.001f;
        s->cell_counts.arr[i] = (int)ceilf(size / part);
        s->cell_sizes.arr[i] = size / (float)s->cell_counts.arr[i];
        cells *= (unsigned long long)s->cell_counts.arr[i];
    }
    if (cells > s->grid_max_cells) {
        float volume_ratio = cells / (float)s->grid_max_cells;
        float ratio = powf(volume_ratio, 1.0f / (float)s->dims);
        for (int i = 0; i < s->dims; ++i) {
            float part = s->min_part_sizes.arr[i];
            float size = s->box.max.arr[i] - s->box.min.arr[i] + part * 0.001f;
            s->cell_counts.arr[i] = (int)ceilf(s->cell_counts.arr[i] / ratio);
            s->cell_sizes.arr[i] = size / (float)s->cell_counts.arr[i];
        }
    }
    s->cells_count = 1;
    for (int i = 0; i < s->dims; ++i)
        s->cells_count *= (unsigned long long)s->cell_counts.arr[i];

    for (unsigned a = 0; a < s->count; ++a) {
        TayAgentTag *agent = (TayAgentTag *)(g->storage + (size_t)g->agent_size * a);
        Int4 idx = cell_indices(POS(agent), s->origin, s->cell_sizes, s->dims);
        unsigned cell_i = (unsigned)row_major(idx, s->cell_counts, s->dims);
        GridCell *cell = s->grid_cells + cell_i;
        if (cell->count == 0) {
            cell->next = s->first_cell;
            s->first_cell = cell;
            cell->indices = idx;
        }
        agent->part_i = cell_i;
        agent->cell_agent_i = cell->count++;
    }

    unsigned first = 0;
    for (GridCell *cell = s->first_cell; cell; cell = cell->next) {
        cell->first = first;
        first += cell->count;
    }

    for (unsigned a = 0; a < s->count; ++a) {
        TayAgentTag *src = (TayAgentTag *)(g->storage + (size_t)g->agent_size * a);
        unsigned dst_i = s->grid_cells[src->part_i].first + src->cell_agent_i;
        memcpy(g->sort_storage + (size_t)g->agent_size * dst_i, src, (size_t)g->agent_size);
    }
    char *t = g->storage;
    g->storage = g->sort_storage;
    g->sort_storage = t;
}

/* tay/cpu_grid.c:195-198 */
static void cpu_grid_unsort(TayGroup *g) {
    for (GridCell *cell = g->space.first_cell; cell; cell = cell->next)
        cell->count = 0;
}

/* Clamp the index range of a seer box to the grid (tay/cpu_grid.c:203-215). */
static void clamp_range(const Space *s, Box box, int dims, Int4 *lo, Int4 *hi) {
    *lo = cell_indices(box.min.arr, s->origin, s->cell_sizes, dims);
    *hi = cell_indices(box.max.arr, s->origin, s->cell_sizes, dims);
    for (int i = 0; i < dims; ++i) {
        if (lo->arr[i] < 0) lo->arr[i] = 0;
        if (hi->arr[i] < 0) hi->arr[i] = 0;
        if (lo->arr[i] >= s->cell_counts.arr[i]) lo->arr[i] = s->cell_counts.arr[i] - 1;
        if (hi->arr[i] >= s->cell_counts.arr[i]) hi->arr[i] = s->cell_counts.arr[i] - 1;
    }
    for (int i = dims; i < 4; ++i)
        lo->arr[i] = hi->arr[i] = 0;
}

typedef void (*CellVisit)(TayPass *p, Slice seer, Int4 idx, PairCounts *c);

/* Visit the cells of [lo, hi] with x outermost, then y, z, w
 * (tay/cpu_grid.c:222-279). */
static void sweep(TayPass *p, Slice seer, Int4 lo, Int4 hi, CellVisit visit, PairCounts *c) {
    Int4 idx = {{0, 0, 0, 0}};
    for (idx.arr[0] = lo.arr[0]; idx.arr[0] <= hi.arr[0]; ++idx.arr[0])
        for (idx.arr[1] = lo.arr[1]; idx.arr[1] <= hi.arr[1]; ++idx.arr[1])
            for (idx.arr[2] = lo.arr[2]; idx.arr[2] <= hi.arr[2]; ++idx.arr[2])
                for (idx.arr[3] = lo.arr[3]; idx.arr[3] <= hi.arr[3]; ++idx.arr[3])
                    visit(p, seer, idx, c);
}

static Slice seen_slice(TayPass *p, unsigned beg, unsigned end) {
    Slice s = {p->seen_group->storage, (unsigned)p->seen_group->agent_size, beg, end};
    return s;
}

static void visit_grid_cell(TayPass *p, Slice seer, Int4 idx, PairCounts *c) {
    const Space *s = &p->seen_group->space;
    GridCell *cell = s->grid_cells + row_major(idx, s->cell_counts, s->dims);
    p->pairing(p, seer, seen_slice(p, cell->first, cell->first + cell->count), c);
}

/* tay/cpu_grid.c:200-280 */
static void cpu_grid_see_seen(TayPass *p, Slice seer, Box box, int dims, PairCounts *c) {
    Int4 lo, hi;
    clamp_range(&p->seen_group->space, box, dims, &lo, &hi);
    sweep(p, seer, lo, hi, visit_grid_cell, c);
}

/* -------------------------------------------------------------- CpuZGrid */

static unsigned floor_log2(unsigned v) {    /* tay/cpu_z_grid.c:128-133 */
    unsigned n = 0;
    while ((1ull << (n + 1u)) <= v)
        ++n;
    return n;
}

static unsigned long long powi(unsigned long long v, int dims) {
    unsigned long long p = 1;
    for (int i = 0; i < dims; ++i)
        p *= v;
    return p;
}

/* tay/cpu_z_grid.c:149-154 */
static void cpu_z_grid_start(Space *s) {
    s->z_cells = (ZCell *)s->shared;
    s->z_max_cells = s->shared_size / CPU_Z_GRID_CELL_BYTES;
    memset(s->z_cells, 0, s->z_max_cells * sizeof(ZCell));
}

/* tay/cpu_z_grid.c:156-221: one power-of-two count for all axes, dense cell
 * array indexed by Morton key, offsets in key order. */
static void cpu_z_grid_sort(TayGroup *g) {
    Space *s = &g->space;
    update_box(g);

    unsigned n = 0;
    for (int i = 0; i < s->dims; ++i) {
        float part = s->min_part_sizes.arr[i];
        float size = s->box.max.arr[i] - s->box.min.arr[i] + part * 0.001f;
        unsigned side_n = floor_log2((unsigned)floorf(size / part));
        if (side_n > n)
            n = side_n;
    }
    while (n > 0 && powi(1ull << (n + 1u), s->dims) > s->z_max_cells)
        --n;

    s->origin = s->box.min;
    for (int i = 0; i < s->dims; ++i) {
        float part = s->min_part_sizes.arr[i];
        float size = s->box.max.arr[i] - s->box.min.arr[i] + part * 0.001f;
        s->cell_counts.arr[i] = 1 << n;
        s->cell_sizes.arr[i] = size / (float)s->cell_counts.arr[i];
    }
    s->cells_count = powi(1ull << n, s->dims);

    for (unsigned a = 0; a < s->count; ++a) {
        TayAgentTag *agent = (TayAgentTag *)(g->storage + (size_t)g->agent_size * a);
        Int4 idx = cell_indices(POS(agent), s->origin, s->cell_sizes, s->dims);
        unsigned cell_i = morton(idx, s->dims);
        agent->part_i = cell_i;
        agent->cell_agent_i = s->z_cells[cell_i].count++;
    }

    unsigned first = 0;
    for (unsigned long long ci = 0; ci < s->cells_count; ++ci) {
        s->z_cells[ci].first = first;
        first += s->z_cells[ci].count;
    }

    for (unsigned a = 0; a < s->count; ++a) {
        TayAgentTag *src = (TayAgentTag *)(g->storage + (size_t)g->agent_size * a);
        unsigned dst_i = s->z_cells[src->part_i].first + src->cell_agent_i;
        memcpy(g->sort_storage + (size_t)g->agent_size * dst_i, src, (size_t)g->agent_size);
    }
    char *t = g->storage;
    g->storage = g->sort_storage;
    g->sort_storage = t;
}

/* tay/cpu_z_grid.c:223-225 */
static void cpu_z_grid_unsort(TayGroup *g) {
    memset(g->space.z_cells, 0, g->space.cells_count * sizeof(ZCell));
}

static void visit_z_cell(TayPass *p, Slice seer, Int4 idx, PairCounts *c) {
    const Space *s = &p->seen_group->space;
    ZCell *cell = s->z_cells + morton(idx, s->dims);
    p->pairing(p, seer, seen_slice(p, cell->first, cell->first + cell->count), c);
}

/* tay/cpu_z_grid.c:232-277: only the 3D sweep exists; 1D, 2D and 4D see
 * nothing. */
static void cpu_z_grid_see_seen(TayPass *p, Slice seer, Box box, int dims, PairCounts *c) {
    if (dims != 3)
        return;
    Int4 lo, hi;
    clamp_range(&p->seen_group->space, box, dims, &lo, &hi);
    sweep(p, seer, lo, hi, visit_z_cell, c);
}

/* ------------------------------------------------- GPU Z-grid statement */
/* Sub-cell bits of the GPU Z-grid's storage order (the GPU library's
 * tay_set_order_bits, same default). */
int tay_set_order_bits(TayState *state, TayGroup *group, int bits) {
    if (bits < 0 || bits > 3) {
        set_error(state, TAY_ERROR_NOT_IMPLEMENTED, "order bits must be 0 to 3");
        return 0;
    }
    group->space.order_bits = bits;
    return 1;
}

/* See order and tile layout of the GPU Z-grid statement (oracle.h). */
int tay_set_see_order(TayState *state, int order) {
    if (order != 0 && order != 1) {
        set_error(state, TAY_ERROR_NOT_IMPLEMENTED, "see order must be TAY_SEE_ORDER_REFERENCE or TAY_SEE_ORDER_BALANCED");
        return 0;
    }
    state->see_order = order;
    return 1;
}

int tay_set_see_split(TayState *state, unsigned pass_index, unsigned lanes, void (*clear)(void *),
                      const char *clear_name, void (*combine)(void *, void *), const char *combine_name) {
    (void)clear_name;
    (void)combine_name;
    if (pass_index >= state->passes_count || state->passes[pass_index].type != PASS_SEE) {
        set_error(state, TAY_ERROR_NOT_IMPLEMENTED, "see split: no see pass at that index");
        return 0;
    }
    if (lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8) {
        set_error(state, TAY_ERROR_NOT_IMPLEMENTED, "see split: lanes per seer must be 1, 2, 4 or 8");
        return 0;
    }
    if (lanes > 1 && (!clear || !combine)) {
        set_error(state, TAY_ERROR_NOT_IMPLEMENTED, "see split: clear and combine functions must be given");
        return 0;
    }
    TayPass *p = state->passes + pass_index;
    p->split_lanes = lanes;
    p->split_clear = clear;
    p->split_combine = combine;
    return 1;
}

int tay_set_tile_seers(TayState *state, unsigned seers) {
    if (seers != 0 && seers != 64 && seers != 32 && seers != 16) {
        set_error(state, TAY_ERROR_NOT_IMPLEMENTED, "seers per tile must be 0, 64, 32 or 16");
        return 0;
    }
    state->tile_seers = seers;
    return 1;
}

int tay_set_tile_width(TayState *state, unsigned width) {
    if (width != 64 && width != 32 && width != 16 && width != 8) {
        set_error(state, TAY_ERROR_NOT_IMPLEMENTED, "tile width must be 64, 32, 16 or 8");
        return 0;
    }
    state->tile_width = width;
    return 1;
}

int tay_set_lane_walk(TayState *state, unsigned width) {
    if (width > 64) {
        set_error(state, TAY_ERROR_NOT_IMPLEMENTED, "lane walk width must be at most 64");
        return 0;
    }
    state->lane_walk = width;
    return 1;
}

/* The geometry is the reference OpenCL Z-grid's (tay/ocl_grid.c:128-136):
 * count_i = ceil(size_i / part_i), cell_size_i = size_i / count_i with
 * size_i = max_i - min_i + 0.001 part_i as in every reference grid.  That
 * kernel has no memory cap; this space coarsens by the smallest integer k with
 * ceil(count_i / k) fitting both the Morton range (1024 per axis in 3D, 65536
 * in 2D) and shared_size / 8 cells.  Agents are ordered by a storage-order
 * key: the cell's Morton key refined by b sub-cell bits per axis (the
 * Morton key of a 2^b-finer grid, b the finest of order_bits that fits the
 * Morton range), ties kept in previous storage order; part_i is the key and cell_agent_i the
 * rank in the cell.  Each seer visits the cells overlapping
 * [p - r - m, p + r + m] (m = (|p| + r) * 2^-20 covers the rounding of
 * p - r and of the box test) x outermost, then y, z, and inside a cell the
 * storage order: the sequence in which the GPU kernel meets each seer's
 * pairs. */

static unsigned morton_axis_limit(int dims) {
    return dims == 3 ? 1024u : dims == 2 ? 65536u : 0xFFFFFFFFu;
}

static void gpu_z_grid_start(Space *s) {
    s->ranges = (CellRange *)s->shared;
    s->gz_max_cells = s->shared_size / GPU_Z_GRID_CELL_BYTES;
}

static void gpu_z_grid_geometry(Space *s) {
    int counts[4] = {1, 1, 1, 1};
    float sizes[4] = {0, 0, 0, 0};
    for (int i = 0; i < s->dims; ++i) {
        float part = s->min_part_sizes.arr[i];
        sizes[i] = s->box.max.arr[i] - s->box.min.arr[i] + part * 0.001f;
        counts[i] = (int)ceilf(sizes[i] / part);
        if (counts[i] < 1)
            counts[i] = 1;
    }
    unsigned limit = morton_axis_limit(s->dims);
    for (unsigned k = 1;; ++k) {
        unsigned long long total = 1;
        int fits = 1;
        for (int i = 0; i < s->dims; ++i) {
            unsigned long long c = ((unsigned long long)counts[i] + k - 1) / k;
            if (c > limit)
                fits = 0;
            total *= c;
        }
        if (fits && total <= s->gz_max_cells) {
            for (int i = 0; i < s->dims; ++i)
                counts[i] = (int)(((unsigned long long)counts[i] + k - 1) / k);
            break;
        }
    }
    s->origin = s->box.min;
    s->cells_count = 1;
    for (int i = 0; i < 4; ++i) {
        s->cell_counts.arr[i] = i < s->dims ? counts[i] : 1;
        s->cell_sizes.arr[i] = i < s->dims ? sizes[i] / (float)counts[i] : 1.0f;
        if (i < s->dims)
            s->cells_count *= (unsigned long long)counts[i];
    }
    /* the finest sub-cell order whose key fits the Morton range */
    int b = s->order_bits;
    for (; b > 0; --b) {
        int fits = 1;
        for (int i = 0; i < s->dims; ++i)
            if (((unsigned long long)counts[i] << b) > limit)
                fits = 0;
        if (fits)
            break;
    }
    s->gz_order_bits = b;
}

static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

static void gpu_z_grid_sort(TayGroup *g) {
    Space *s = &g->space;
    update_box(g);
    gpu_z_grid_geometry(s);
    /* non-point agents are binned by their min corner; a seer's seen range
     * then reaches down by the largest extent of the group */
    for (int k = 0; k < 4; ++k)
        s->gz_max_extent[k] = 0.0f;
    if (!g->is_point)
        for (unsigned a = 0; a < s->count; ++a) {
            const char *agent = g->storage + (size_t)g->agent_size * a;
            for (int k = 0; k < s->dims; ++k) {
                float e = MAXC(agent)[k] - POS(agent)[k];
                if (e > s->gz_max_extent[k])
                    s->gz_max_extent[k] = e;
            }
        }

    /* Storage-order key: per axis q = (p - o) / cs, the cell c = floorf(q)
     * (clamped: only a fixed box can leave agents outside) refined by
     * b sub-cell bits, sub = floorf((q - c) * 2^b) clamped to [0, 2^b - 1];
     * the key is the Morton interleave of (c << b) | sub, whose top bits are
     * the cell's Morton key.  Ties keep the previous storage order. */
    const int b = s->gz_order_bits;
    const unsigned shift = (unsigned)(s->dims * b);
    s->sort_buf = (uint64_t *)realloc(s->sort_buf, sizeof(uint64_t) * (s->count ? s->count : 1));
    for (unsigned a = 0; a < s->count; ++a) {
        const float *p = POS(g->storage + (size_t)g->agent_size * a);
        Int4 idx = {{0, 0, 0, 0}};
        for (int k = 0; k < s->dims; ++k) {
            float q = (p[k] - s->origin.arr[k]) / s->cell_sizes.arr[k];
            int c = (int)floorf(q);
            if (c < 0) c = 0;
            if (c >= s->cell_counts.arr[k]) c = s->cell_counts.arr[k] - 1;
            int sub = 0;
            if (b) {
                sub = (int)floorf((q - (float)c) * (float)(1 << b));
                if (sub < 0) sub = 0;
                if (sub > (1 << b) - 1) sub = (1 << b) - 1;
            }
            idx.arr[k] = (c << b) | sub;
        }
        s->sort_buf[a] = ((uint64_t)morton(idx, s->dims) << 32) | a;
    }
    qsort(s->sort_buf, s->count, sizeof(uint64_t), cmp_u64);

    for (unsigned long long ci = 0; ci < s->cells_count; ++ci)
        s->ranges[ci].beg = s->ranges[ci].end = 0;
    unsigned run_beg = 0;
    for (unsigned i = 0; i < s->count; ++i) {
        unsigned key = (unsigned)(s->sort_buf[i] >> 32) >> shift;
        unsigned src = (unsigned)(s->sort_buf[i] & 0xFFFFFFFFu);
        if (i > 0 && key != (unsigned)(s->sort_buf[i - 1] >> 32) >> shift)
            run_beg = i;
        TayAgentTag *dst = (TayAgentTag *)(g->sort_storage + (size_t)g->agent_size * i);
        memcpy(dst, g->storage + (size_t)g->agent_size * src, (size_t)g->agent_size);
        dst->part_i = key;
        dst->cell_agent_i = i - run_beg;
        if (i + 1 == s->count || key != (unsigned)(s->sort_buf[i + 1] >> 32) >> shift) {
            int x, y, z;
            oracle_morton_decode(key, s->dims, &x, &y, &z);
            Int4 idx = {{x, y, z, 0}};
            CellRange *r = s->ranges + row_major(idx, s->cell_counts, s->dims);
            r->beg = run_beg;
            r->end = i + 1;
        }
    }
    char *t = g->storage;
    g->storage = g->sort_storage;
    g->sort_storage = t;
}

static void visit_gpu_z_cell(TayPass *p, Slice seer, Int4 idx, PairCounts *c) {
    const Space *s = &p->seen_group->space;
    CellRange *r = s->ranges + row_major(idx, s->cell_counts, s->dims);
    p->pairing(p, seer, seen_slice(p, r->beg, r->end), c);
}

/* Per-seer conservative cell range: the seen min corners (positions) that
 * can pass the box test lie in [lo - r - extent, hi + r], extent the seen
 * group's largest box extent (0 for points), widened by a margin of
 * (|p| + r + extent) * 2^-20 for the rounding of these sums and of the test. */
static void gpu_z_grid_range(const Space *s, const float *lo_pos, const float *hi_pos, const float *r, int dims, Int4 *lo, Int4 *hi) {
    Box box;
    for (int i = 0; i < dims; ++i) {
        float ext = s->gz_max_extent[i];
        float mlo = (fabsf(lo_pos[i]) + r[i] + ext) * 0x1p-20f;
        float mhi = (fabsf(hi_pos[i]) + r[i]) * 0x1p-20f;
        box.min.arr[i] = ((lo_pos[i] - r[i]) - ext) - mlo;
        box.max.arr[i] = (hi_pos[i] + r[i]) + mhi;
    }
    clamp_range(s, box, dims, lo, hi);
}

static void gpu_z_grid_see_seen_one(TayPass *p, Slice seer, PairCounts *c) {
    const Space *s = &p->seen_group->space;
    int dims = s->dims;
    for (unsigned i = seer.beg; i < seer.end; ++i) {
        const char *a = AGENT_AT(seer, i);
        const float *lo_pos = POS(a);
        const float *hi_pos = p->seer_group->is_point ? lo_pos : MAXC(a);
        Int4 lo, hi;
        gpu_z_grid_range(s, lo_pos, hi_pos, p->radii.arr, dims, &lo, &hi);
        Slice one = seer;
        one.beg = i;
        one.end = i + 1;
        sweep(p, one, lo, hi, visit_gpu_z_cell, c);
    }
}

/* Balanced see order: the GPU library's see kernel (jit_device.h,
 * tay_see_tile) takes the seers of a seer group in tiles of S consecutive
 * storage slots and each tile in lane groups: a group starts `width` lanes
 * wide (the tile width, at most S); when its box grown by the radii spans
 * more than `limit` cells it halves while wider than the lane walk, else
 * every member sweeps its own cells in the reference order (the per-lane
 * walk); otherwise the group sweeps the cells of its box x, then y, then z,
 * 64 cells at a time.  The M candidates of such a segment, in sweep order,
 * are cut into P = 64 / BALANCE_RUN parts of L = ceil(M / P); chunk k holds
 * run k (BALANCE_RUN consecutive positions) of every part, slot
 * j * BALANCE_RUN + t holding position j * L + k * BALANCE_RUN + t when that
 * lies inside part j and the segment, and every member meets its passing
 * candidates chunk by chunk, slot by slot.  After a group the width doubles
 * back while the next start is aligned to it.  Float expressions follow the
 * kernel's. */
#define BALANCE_RUN 4u      /* the GPU library's TAY_BALANCE_RUN */

static Slice seer_slice_of(TayPass *p, unsigned beg, unsigned end);
static Slice seen_slice(TayPass *p, unsigned beg, unsigned end);

/* seers per tile chosen by the seer count: 64 while that gives each of a
 * device's SIMDs two tiles, else 32, else 16 (jit.cpp, auto_tile_seers) */
static unsigned auto_tile_seers(unsigned seers) {
    const unsigned simds = 1024u;   /* 256 compute units of 4 SIMDs */
    for (unsigned s = 64u; s > 16u; s >>= 1)
        if ((seers + s - 1u) / s >= 2u * simds)
            return s;
    return 16u;
}

static int gpu_cell_axis(float p, float origin, float size, int count) {
    int c = (int)floorf((p - origin) / size);
    c = c < 0 ? 0 : c;
    return c > count - 1 ? count - 1 : c;
}

static void gpu_z_grid_balanced_sweep(TayPass *p, unsigned m0, unsigned m1, const int *lo, const int *hi,
                                      PairCounts *c) {
    const Space *s = &p->seen_group->space;
    Slice seers = seer_slice_of(p, 0, 0);
    const unsigned nz = (unsigned)(hi[2] - lo[2] + 1), nyz = nz * (unsigned)(hi[1] - lo[1] + 1);
    const unsigned swept = nyz * (unsigned)(hi[0] - lo[0] + 1);
    for (unsigned c0 = 0; c0 < swept; c0 += 64u) {
        const unsigned n = swept - c0 < 64u ? swept - c0 : 64u;
        unsigned cbeg[64], pre[64], total = 0;
        for (unsigned q = 0; q < n; ++q) {
            const unsigned cell = c0 + q, rem = cell % nyz;
            Int4 idx = {{lo[0] + (int)(cell / nyz), lo[1] + (int)(rem / nz), lo[2] + (int)(rem % nz), 0}};
            const CellRange *r = s->ranges + row_major(idx, s->cell_counts, s->dims);
            cbeg[q] = r->beg;
            pre[q] = total;
            total += r->end - r->beg;
        }
        if (!total)
            continue;
        const unsigned parts = 64u / BALANCE_RUN;
        const unsigned part_len = (total + parts - 1u) / parts;
        const unsigned chunks = (part_len + BALANCE_RUN - 1u) / BALANCE_RUN;
        for (unsigned k = 0; k < chunks; ++k)
            for (unsigned slot = 0; slot < 64u; ++slot) {
                const unsigned o = k * BALANCE_RUN + slot % BALANCE_RUN;
                const unsigned pos = slot / BALANCE_RUN * part_len + o;
                if (o >= part_len || pos >= total)
                    continue;
                unsigned q = 0;     /* the last cell whose prefix is at most pos */
                while (q + 1u < n && pre[q + 1u] <= pos)
                    ++q;
                const unsigned j = cbeg[q] + (pos - pre[q]);
                for (unsigned i = m0; i < m1; ++i) {
                    Slice one = seers;
                    one.beg = i;
                    one.end = i + 1;
                    p->pairing(p, one, seen_slice(p, j, j + 1), c);
                }
            }
    }
}

/* One pair for a seer record that may be a partial (split pass): seer index
 * i (for the self test) and seen index j, with the box test of the groups'
 * agent kinds and the self rule of the pairing loops (tay/space.c:106-238). */
static void pair_one(TayPass *p, char *rec, unsigned i, unsigned j, PairCounts *c) {
    if (p->seer_group == p->seen_group && !p->self_see && i == j)
        return;
    const char *b = p->seen_group->storage + (size_t)p->seen_group->agent_size * j;
    const int dims = p->seer_group->space.dims;
    const float *r = p->radii.arr;
    const int sp = p->seer_group->is_point, np = p->seen_group->is_point;
    ++c->broad;
    int ok;
    if (sp)
        ok = np ? narrow_point_point(POS(rec), POS(b), r, dims) : narrow_point_nonpoint(POS(rec), POS(b), MAXC(b), r, dims);
    else
        ok = np ? narrow_nonpoint_point(POS(rec), MAXC(rec), POS(b), r, dims)
                : narrow_nonpoint_nonpoint(POS(rec), MAXC(rec), POS(b), MAXC(b), r, dims);
    if (!ok)
        return;
    ++c->narrow;
    p->see(rec, (void *)b, p->context);
}

/* Split see pass (tay_set_see_split, K = split_lanes): the kernel's lanes
 * of a seer share its lane group's sweep, candidates at sweep positions
 * k (mod K) going to partial k; rec[m * K + k] is member m's partial k
 * (partial 0 the seer's own record). */
static void gpu_z_grid_split_sweep(TayPass *p, unsigned m0, unsigned m1, const int *lo, const int *hi, char **rec,
                                   PairCounts *c) {
    const Space *s = &p->seen_group->space;
    const unsigned K = p->split_lanes;
    unsigned pos = 0;
    Int4 idx = {{0, 0, 0, 0}};
    for (idx.arr[0] = lo[0]; idx.arr[0] <= hi[0]; ++idx.arr[0])
        for (idx.arr[1] = lo[1]; idx.arr[1] <= hi[1]; ++idx.arr[1])
            for (idx.arr[2] = lo[2]; idx.arr[2] <= hi[2]; ++idx.arr[2]) {
                const CellRange *cr = s->ranges + row_major(idx, s->cell_counts, s->dims);
                for (unsigned j = cr->beg; j < cr->end; ++j, ++pos)
                    for (unsigned i = m0; i < m1; ++i)
                        pair_one(p, rec[(i - m0) * K + pos % K], i, j, c);
            }
}

/* The per-seer walk of a split pass: each member walks the cells of its own
 * box (gpu_z_grid_range) in x, y, z order, the cells of row (x, y) going to
 * partial k when the row's index in the box, (x - lo.x) * ny + (y - lo.y),
 * is k mod K. */
static void gpu_z_grid_split_walk(TayPass *p, unsigned m0, unsigned m1, char **rec, PairCounts *c) {
    const Space *s = &p->seen_group->space;
    const unsigned K = p->split_lanes;
    for (unsigned i = m0; i < m1; ++i) {
        const char *a = p->seer_group->storage + (size_t)p->seer_group->agent_size * i;
        const float *lo_pos = POS(a);
        const float *hi_pos = p->seer_group->is_point ? lo_pos : MAXC(a);
        Int4 lo, hi, idx = {{0, 0, 0, 0}};
        gpu_z_grid_range(s, lo_pos, hi_pos, p->radii.arr, s->dims, &lo, &hi);
        const unsigned ny = (unsigned)(hi.arr[1] - lo.arr[1] + 1);
        for (idx.arr[0] = lo.arr[0]; idx.arr[0] <= hi.arr[0]; ++idx.arr[0])
            for (idx.arr[1] = lo.arr[1]; idx.arr[1] <= hi.arr[1]; ++idx.arr[1]) {
                const unsigned row = (unsigned)(idx.arr[0] - lo.arr[0]) * ny + (unsigned)(idx.arr[1] - lo.arr[1]);
                char *r = rec[(i - m0) * K + row % K];
                for (idx.arr[2] = lo.arr[2]; idx.arr[2] <= hi.arr[2]; ++idx.arr[2]) {
                    const CellRange *cr = s->ranges + row_major(idx, s->cell_counts, s->dims);
                    for (unsigned j = cr->beg; j < cr->end; ++j)
                        pair_one(p, r, i, j, c);
                }
            }
    }
}

/* Seers [beg, end) of the seer group, whole tiles of S, out of `seers`. */
static void gpu_z_grid_see_tiles(TayPass *p, unsigned beg, unsigned end, unsigned seers, unsigned S,
                                 PairCounts *c) {
    const Space *s = &p->seen_group->space;
    const TayGroup *sg = p->seer_group;
    const int dims = s->dims;
    const float *r = p->radii.arr;
    float ext[3] = {0.0f, 0.0f, 0.0f};
    unsigned limit = 1;
    for (int k = 0; k < dims; ++k) {
        if (!p->seen_group->is_point)
            ext[k] = s->gz_max_extent[k];
        limit *= (unsigned)ceilf((2.0f * r[k] + ext[k]) / s->cell_sizes.arr[k]) + 3u;
    }
    const unsigned width0 = p->tile_width < S ? p->tile_width : S;
    /* a split pass: partial records 1..K-1 of the tile's seers */
    const unsigned K = p->split_lanes > 1u ? p->split_lanes : 1u;
    const size_t asize = (size_t)sg->agent_size;
    char *rec[64];          /* a split pass has S * K = 64 partials per tile */
    char *parts = NULL;
    if (K > 1u) {
        parts = (char *)malloc(asize * S * (K - 1u));
        if (!parts) {
            fprintf(stderr, "oracle: no memory for the split pass's partial records\n");
            abort();
        }
    }
    for (unsigned i0 = beg; i0 < end; i0 += S) {
        const unsigned n_active = seers - i0 < S ? seers - i0 : S;
        if (K > 1u)
            for (unsigned m = 0; m < n_active; ++m) {
                rec[m * K] = sg->storage + asize * (i0 + m);
                for (unsigned k = 1; k < K; ++k) {
                    char *q = parts + asize * (m * (K - 1u) + (k - 1u));
                    memcpy(q, rec[m * K], asize);
                    p->split_clear(q);
                    rec[m * K + k] = q;
                }
            }
        unsigned gbeg = 0, width = width0;
        while (gbeg < n_active) {
            const unsigned m0 = i0 + gbeg, m1 = i0 + (gbeg + width < n_active ? gbeg + width : n_active);
            int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
            unsigned cells = 1;
            for (int k = 0; k < dims; ++k) {
                float tmin = INFINITY, tmax = -INFINITY;
                for (unsigned i = m0; i < m1; ++i) {
                    const char *a = sg->storage + (size_t)sg->agent_size * i;
                    tmin = fminf(tmin, POS(a)[k]);
                    tmax = fmaxf(tmax, sg->is_point ? POS(a)[k] : MAXC(a)[k]);
                }
                const float mlo = (fabsf(tmin) + r[k] + ext[k]) * 0x1p-20f;
                const float mhi = (fabsf(tmax) + r[k]) * 0x1p-20f;
                lo[k] = gpu_cell_axis(((tmin - r[k]) - ext[k]) - mlo, s->origin.arr[k], s->cell_sizes.arr[k],
                                      s->cell_counts.arr[k]);
                hi[k] = gpu_cell_axis((tmax + r[k]) + mhi, s->origin.arr[k], s->cell_sizes.arr[k],
                                      s->cell_counts.arr[k]);
                cells *= (unsigned)(hi[k] - lo[k] + 1);
            }
            if (cells > limit && width > 1u && width > p->lane_walk) {
                width >>= 1;
                continue;
            }
            if (cells > limit && width > 1u) {
                if (K > 1u)
                    gpu_z_grid_split_walk(p, m0, m1, rec + (m0 - i0) * K, c);
                else
                    gpu_z_grid_see_seen_one(p, seer_slice_of(p, m0, m1), c);
            } else if (hi[0] >= lo[0] && hi[1] >= lo[1] && hi[2] >= lo[2]) {
                if (K > 1u)
                    gpu_z_grid_split_sweep(p, m0, m1, lo, hi, rec + (m0 - i0) * K, c);
                else
                    gpu_z_grid_balanced_sweep(p, m0, m1, lo, hi, c);
            }
            gbeg += width;
            while (width < width0 && (gbeg & (2u * width - 1u)) == 0u)
                width <<= 1;
        }
        /* merge the partials in the kernel's tree: (0 + 1), (2 + 3), ...,
         * then pairs of those */
        if (K > 1u)
            for (unsigned m = 0; m < n_active; ++m)
                for (unsigned step = 1; step < K; step <<= 1)
                    for (unsigned k = 0; k < K; k += 2u * step)
                        p->split_combine(rec[m * K + k], rec[m * K + k + step]);
    }
    free(parts);
}

/* -------------------------------------------------------------- see/act */

typedef struct {
    TayPass *pass;
} PassTask;

static Slice seer_slice_of(TayPass *p, unsigned beg, unsigned end) {
    Slice s = {p->seer_group->storage, (unsigned)p->seer_group->agent_size, beg, end};
    return s;
}

/* Seen side dispatch by the seen space's type (tay/state_compile.c:122-135). */
static void see_seen(TayPass *p, Slice seer, Box box, int dims, PairCounts *c) {
    switch (p->seen_group->space.type) {
        case TAY_CPU_GRID: cpu_grid_see_seen(p, seer, box, dims, c); break;
        case TAY_CPU_Z_GRID: cpu_z_grid_see_seen(p, seer, box, dims, c); break;
        case TAY_GPU_Z_GRID: gpu_z_grid_see_seen_one(p, seer, c); break;
        default: /* CpuSimple: every seen agent (tay/cpu_simple.c:6-16) */
            p->pairing(p, seer, seen_slice(p, 0, p->seen_group->space.count), c);
    }
}

/* Seer drivers (tay/cpu_simple.c:18-38, tay/cpu_grid.c:286-323,
 * tay/cpu_z_grid.c:283-321): the thread's seer range is walked one cell run
 * at a time with the cell's box grown by the radii as the seer box. */
static void see_task(void *task_p, int thread_i, PairCounts *c) {
    TayPass *p = ((PassTask *)task_p)->pass;
    TayGroup *seer = p->seer_group;
    const Space *ss = &seer->space;
    int dims = ss->dims < p->seen_group->space.dims ? ss->dims : p->seen_group->space.dims;
    unsigned beg, end;
    unsigned seers = p->seer_limit && p->seer_limit < ss->count ? p->seer_limit : ss->count;
    const int split = p->split_lanes > 1u;
    if ((p->see_order == 1 || split) && p->seen_group->space.type == TAY_GPU_Z_GRID && ss->type != TAY_CPU_GRID &&
        ss->type != TAY_CPU_Z_GRID) {
        /* balanced order or split pass: each thread takes whole tiles; a
         * split pass has 64 / lanes seers per tile and the sweep order */
        const unsigned S = split ? 64u / p->split_lanes : p->tile_seers ? p->tile_seers : auto_tile_seers(ss->count);
        unsigned tb, te;
        thread_range((seers + S - 1u) / S, thread_i, &tb, &te);
        gpu_z_grid_see_tiles(p, tb * S, te * S < seers ? te * S : seers, seers, S, c);
        return;
    }
    thread_range(seers, thread_i, &beg, &end);

    if (ss->type != TAY_CPU_GRID && ss->type != TAY_CPU_Z_GRID) {
        TaySpaceType seen_type = p->seen_group->space.type;
        if (seen_type != TAY_CPU_GRID && seen_type != TAY_CPU_Z_GRID) {
            Box none;
            memset(&none, 0, sizeof(none));
            see_seen(p, seer_slice_of(p, beg, end), none, dims, c);
            return;
        }
        /* The reference hands an all-zero box to the seen side here
         * (tay/cpu_simple.c:37), which only suits a Simple seen space; the
         * oracle gives each seer its own box, grown by the radii. */
        for (unsigned i = beg; i < end; ++i) {
            const char *a = seer->storage + (size_t)seer->agent_size * i;
            const float *lo = POS(a);
            const float *hi = seer->is_point ? lo : MAXC(a);
            Box box;
            for (int k = 0; k < dims; ++k) {
                box.min.arr[k] = lo[k] - p->radii.arr[k];
                box.max.arr[k] = hi[k] + p->radii.arr[k];
            }
            see_seen(p, seer_slice_of(p, i, i + 1), box, dims, c);
        }
        return;
    }

    unsigned i = beg;
    while (i < end) {
        TayAgentTag *a = (TayAgentTag *)(seer->storage + (size_t)seer->agent_size * i);
        Int4 idx;
        unsigned cell_first, cell_count;
        if (ss->type == TAY_CPU_GRID) {
            GridCell *cell = ss->grid_cells + a->part_i;
            idx = cell->indices;
            cell_first = cell->first;
            cell_count = cell->count;
        } else {
            ZCell *cell = ss->z_cells + a->part_i;
            int x, y, z;
            oracle_morton_decode(a->part_i, dims, &x, &y, &z);
            idx.arr[0] = x; idx.arr[1] = y; idx.arr[2] = z; idx.arr[3] = 0;
            cell_first = cell->first;
            cell_count = cell->count;
        }
        Box box;
        for (int k = 0; k < dims; ++k) {
            float lo = ss->origin.arr[k] + idx.arr[k] * ss->cell_sizes.arr[k];
            box.min.arr[k] = lo - p->radii.arr[k];
            box.max.arr[k] = lo + ss->cell_sizes.arr[k] + p->radii.arr[k];
        }
        unsigned cell_end = cell_first + cell_count;
        if (cell_end > end)
            cell_end = end;
        see_seen(p, seer_slice_of(p, i, cell_end), box, dims, c);
        i = cell_end;
    }
}

/* tay/space.c:250-263 */
static void act_task(void *task_p, int thread_i, PairCounts *c) {
    (void)c;
    TayPass *p = ((PassTask *)task_p)->pass;
    TayGroup *g = p->seer_group;
    unsigned beg, end;
    thread_range(g->space.count, thread_i, &beg, &end);
    for (unsigned i = beg; i < end; ++i)
        p->act(g->storage + (size_t)g->agent_size * i, p->context);
}

/* --------------------------------------------------------------- compile */

static int is_grid(TaySpaceType t) {
    return t == TAY_CPU_GRID || t == TAY_CPU_Z_GRID || t == TAY_GPU_Z_GRID;
}

/* tay/state_compile.c:5-175 (CPU branch) */
static int state_compile(TayState *state) {
    if (!state->recompile)
        return 1;
    state->recompile = 0;

    for (int gi = 0; gi < MAX_GROUPS; ++gi) {
        TayGroup *g = state->groups + gi;
        if (!g->active)
            continue;
        TaySpaceType t = g->space.type;
        if (t == TAY_CPU_KD_TREE || t == TAY_CPU_AABB_TREE || t == TAY_GPU_SIMPLE) {
            set_error(state, TAY_ERROR_NOT_IMPLEMENTED, "space type not restated in the oracle");
            return 0;
        }
        /* the reference's grids take points only (tay/state_compile.c:25-27);
         * the GPU Z-grid also bins boxes, by their min corner */
        if (!g->is_point && is_grid(t) && t != TAY_GPU_Z_GRID) {
            set_error(state, TAY_ERROR_POINT_NONPOINT_MISMATCH, "grid structure can only contain point agents");
            return 0;
        }
        if (t == TAY_GPU_Z_GRID && (g->space.dims < 1 || g->space.dims > 3)) {
            set_error(state, TAY_ERROR_NOT_IMPLEMENTED, "z-grid supports 1, 2 or 3 dimensions");
            return 0;
        }
    }

    for (unsigned pi = 0; pi < state->passes_count; ++pi) {
        TayPass *p = state->passes + pi;
        if (p->type == PASS_SEE) {
            if (p->see == 0) {
                set_error(state, TAY_ERROR_OCL, "cpu pass with no see func set");
                return 0;
            }
            if (p->seer_group->space.dims != p->seen_group->space.dims) {
                set_error(state, TAY_ERROR_NOT_IMPLEMENTED, "groups with different numbers of dimensions not supported");
                return 0;
            }
            int sp = p->seer_group->is_point, np = p->seen_group->is_point;
            if (p->seer_group == p->seen_group)
                p->pairing = p->self_see ? (sp ? pair_point_point_self_see : pair_nonpoint_nonpoint_self_see)
                                         : (sp ? pair_point_point : pair_nonpoint_nonpoint);
            else if (sp == np)
                p->pairing = sp ? pair_point_point_self_see : pair_nonpoint_nonpoint_self_see;
            else
                p->pairing = sp ? pair_point_nonpoint : pair_nonpoint_point;
        } else if (p->act == 0) {
            set_error(state, TAY_ERROR_OCL, "cpu pass with no act func set");
            return 0;
        }
    }

    for (int gi = 0; gi < MAX_GROUPS; ++gi) {
        TayGroup *g = state->groups + gi;
        if (!g->active)
            continue;
        if (is_grid(g->space.type) && g->space.shared == 0) {
            set_error(state, TAY_ERROR_STATE_STATUS, "grid space has no shared memory configured");
            return 0;
        }
        if (g->space.type == TAY_CPU_GRID)
            cpu_grid_start(&g->space);
        else if (g->space.type == TAY_CPU_Z_GRID)
            cpu_z_grid_start(&g->space);
        else if (g->space.type == TAY_GPU_Z_GRID)
            gpu_z_grid_start(&g->space);
    }
    return 1;
}

/* ------------------------------------------------------------------- run */

static double now_seconds(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + (double)t.tv_nsec * 1e-9;
}

/* tay/state.c:194-323: per step sort every group, run the passes in
 * registration order, unsort every group. */
int tay_run(TayState *state, int steps) {
    if (!state_compile(state))
        return 0;
    if (!state->is_running) {
        set_error(state, TAY_ERROR_STATE_STATUS, "cannot run simulation steps if simulation is not started");
        return 0;
    }
    for (unsigned pi = 0; pi < state->passes_count; ++pi)
        state->passes[pi].counts.broad = state->passes[pi].counts.narrow = 0;
    state->see_seconds = 0.0;

    double beg = now_seconds();
    for (int step = 0; step < steps; ++step) {
        for (int gi = 0; gi < MAX_GROUPS; ++gi) {
            TayGroup *g = state->groups + gi;
            if (!g->active)
                continue;
            if (g->space.type == TAY_CPU_GRID)
                cpu_grid_sort(g);
            else if (g->space.type == TAY_CPU_Z_GRID)
                cpu_z_grid_sort(g);
            else if (g->space.type == TAY_GPU_Z_GRID)
                gpu_z_grid_sort(g);
        }
        for (unsigned pi = 0; pi < state->passes_count; ++pi) {
            TayPass *p = state->passes + pi;
            PassTask task = {p};
            if (p->type == PASS_SEE) {
                p->see_order = state->see_order;
                p->tile_seers = state->tile_seers;
                p->tile_width = state->tile_width;
                p->lane_walk = state->lane_walk;
                double t0 = now_seconds();
                run_tasks(see_task, &task, &p->counts);
                state->see_seconds += now_seconds() - t0;
            } else {
                run_tasks(act_task, &task, 0);
            }
        }
        for (int gi = 0; gi < MAX_GROUPS; ++gi) {
            TayGroup *g = state->groups + gi;
            if (!g->active)
                continue;
            if (g->space.type == TAY_CPU_GRID)
                cpu_grid_unsort(g);
            else if (g->space.type == TAY_CPU_Z_GRID)
                cpu_z_grid_unsort(g);
        }
    }
    double t = now_seconds() - beg;
    state->ms_per_step = steps > 0 ? t / (double)steps * 1000.0 : 0.0;
    return steps;
}

/* ------------------------------------------------------------ inspection */

void oracle_get_pair_counts(TayState *state, unsigned pass_index, unsigned long long *broad, unsigned long long *narrow) {
    if (pass_index >= state->passes_count) {
        *broad = *narrow = 0;
        return;
    }
    *broad = state->passes[pass_index].counts.broad;
    *narrow = state->passes[pass_index].counts.narrow;
}

void oracle_get_grid_info(TayState *state, TayGroup *group, float4 *origin, float4 *cell_sizes, int *cell_counts, unsigned long long *cells_count) {
    (void)state;
    *origin = group->space.origin;
    *cell_sizes = group->space.cell_sizes;
    for (int i = 0; i < 4; ++i)
        cell_counts[i] = group->space.cell_counts.arr[i];
    *cells_count = group->space.cells_count;
}

double oracle_get_see_seconds(TayState *state) {
    return state->see_seconds;
}

void oracle_limit_seers(TayState *state, unsigned pass_index, unsigned count) {
    if (pass_index < state->passes_count)
        state->passes[pass_index].seer_limit = count;
}
