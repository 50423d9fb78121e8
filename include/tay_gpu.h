/*
 * tay_gpu.h - C-ABI of the MI355X GPU spaces (libtay_amd.so).
 *
 * These entry points replace the reference's OpenCL backend at the points
 * where its state machine dispatches to the device (tay/state.c:209-218,
 * 230-233, 253-254, 273-274, 296-299, 313-315; declarations in
 * tay/ocl.h:43-96).  They take plain pointers and sizes only, so a host that
 * keeps its own state machine can call them directly; this library's own
 * tay.h implementation is built on exactly these calls.
 *
 * All device work is queued on one HIP stream owned by the library and runs
 * in submission order.  Every call returns 1 on success and 0 on failure;
 * tay_gpu_last_error() then describes the failure.  Agents live on the device
 * as structure-of-arrays: one column of 32-bit words per word of the agent
 * record.
 */
#ifndef tay_gpu_h
#define tay_gpu_h

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct TayGpuSpace TayGpuSpace;
typedef struct TayGpuProgram TayGpuProgram;

typedef enum TayGpuSpaceKind {
    TAY_GPU_KIND_Z_GRID = 1,    /* Morton-ordered uniform grid, point agents */
    TAY_GPU_KIND_SIMPLE = 2,    /* unsorted, every seer meets every seen agent */
} TayGpuSpaceKind;

typedef enum TayGpuPassType {
    TAY_GPU_PASS_SEE = 0,
    TAY_GPU_PASS_ACT = 1,
} TayGpuPassType;

/* Reports which device the library uses; 0 when no GPU is present.
 * Replaces ocl_init_context (tay/ocl_wrappers.c:213-345). */
int tay_gpu_available(void);

/* Message of the last failed call on this thread ("" if none). */
const char *tay_gpu_last_error(void);

/* Waits for all queued device work.  Replaces ocl_finish (tay/ocl.h:94). */
int tay_gpu_synchronize(void);

/* Zeroed page-locked host memory of `bytes` (at most 4 GiB), for agent
 * storage that tay_gpu_space_push / tay_gpu_space_fetch copy at the link's
 * full rate; NULL without a GPU or when the allocation fails, and the
 * caller then uses ordinary memory.  Replaces the calloc of a group's
 * storage (tay/state.c:88-89).  Free with tay_gpu_host_free. */
void *tay_gpu_host_alloc(size_t bytes);
void tay_gpu_host_free(void *p);

/* ------------------------------------------------------------- spaces */

/* Device storage for one agent group and its space.  agent_size must be a
 * multiple of 4.  For a Z-grid, part_sizes[0..dims) are the suggested cell
 * sizes and shared_bytes bounds the cell table (8 bytes per cell), as the
 * reference's shared_size_in_megabytes bounds its grids
 * (tay/state.c:109-118).  Replaces the per-group part of ocl_create_buffers
 * (tay/ocl.c:59-98). */
TayGpuSpace *tay_gpu_space_create(int kind, unsigned agent_size, unsigned capacity, int is_point,
                                  int dims, const float *part_sizes, unsigned long long shared_bytes);
void tay_gpu_space_destroy(TayGpuSpace *space);

/* Uses [min, max] as the grid box instead of the agents' bounding box; agents
 * outside are clamped into the border cells.  Replaces the host-only
 * tay_fix_space_box (tay/state.c:120-125), which the OpenCL grid ignores
 * (todo:6). */
int tay_gpu_space_fix_box(TayGpuSpace *space, const float *min, const float *max);
int tay_gpu_space_unfix_box(TayGpuSpace *space);

/* Sub-cell bits (0 to 3, default 2) of the Z-grid's storage order inside a
 * cell: agents of a cell are ordered by the Morton key of a 2^bits-finer
 * grid, then by previous storage order, so the 64 seers of a wavefront and
 * each 64-candidate chunk are spatially compact.  0 keeps the previous
 * storage order alone (the reference's CPU grids, tay/cpu_grid.c:114-150).
 * The reference's OpenCL grid ranks agents in a cell by atomic_inc
 * (tay/ocl_grid.c:147): no fixed order at all.  No reference counterpart as
 * a setting. */
int tay_gpu_space_set_order_bits(TayGpuSpace *space, int bits);

/* Copies count AoS records from host memory to the device and makes them the
 * space's agents; the host records may change once it returns.  Replaces
 * ocl_push_agents_non_blocking (tay/ocl.c:238-250). */
int tay_gpu_space_push(TayGpuSpace *space, const void *agents, unsigned count);

/* Copies the space's agents, in current storage order, to count AoS records
 * in host memory.  Replaces ocl_fetch_agents (tay/ocl.c:261-272). */
int tay_gpu_space_fetch(TayGpuSpace *space, void *agents, unsigned count);

/* Copies words [first_word, first_word + words) of the space's first count
 * agents, in current storage order, to host memory as count packed records
 * of 4 * words bytes: the selected fields only (positions for drawing, say),
 * instead of whole agents.  Replaces the whole-agent read of
 * ocl_fetch_agents (tay/ocl.c:261-272) where the caller needs a few fields
 * (the reference's viewer draws positions after every step,
 * entorama/main.c:192-196). */
int tay_gpu_space_fetch_words(TayGpuSpace *space, unsigned first_word, unsigned words, void *dst, unsigned count);

/* The same copy without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight viewer loops, harnesses that forbid hipMalloc):
 * tay_gpu_space_fetch_words allocates its staging buffer on first use and on
 * growth, this variant stages in a fixed 32 MiB device array that each GPU
 * holds until the process ends (count x words x 4 bytes must fit it, or the
 * call fails with an error naming both sizes).  Every call on a GPU shares
 * that array: calls that overlap on one GPU (from several host threads)
 * race.  For a copy without host-side staging dst should be page-locked
 * (tay_gpu_host_alloc). */
int tay_gpu_space_fetch_words_static(TayGpuSpace *space, unsigned first_word, unsigned words, void *dst,
                                     unsigned count);

unsigned tay_gpu_space_count(const TayGpuSpace *space);

/* Rebuilds a Z-grid from the current positions: bounding box, geometry, Morton
 * keys, stable radix sort, cell ranges, column permutation.  No host
 * synchronisation.  A no-op for a Simple space.  Replaces
 * ocl_grid_run_sort_kernel (tay/ocl_grid.c:241-362); the cell reset of
 * ocl_grid_run_unsort_kernel (tay/ocl_grid.c:364-374) has no counterpart
 * because cell ranges are rebuilt, not accumulated. */
int tay_gpu_space_sort(TayGpuSpace *space);

/* Geometry of the last sort (synchronises).  cell_counts has 4 entries. */
int tay_gpu_space_grid_info(TayGpuSpace *space, float *origin, float *cell_sizes, int *cell_counts,
                            unsigned long long *cells_count);

/* Diagnostic: one streaming copy of `bytes` between two fresh device buffers
 * (reads and writes `bytes`), timed with HIP events.  The calibration kernel
 * for HBM traffic counters and a measured copy bandwidth; no reference
 * counterpart. */
int tay_gpu_copy_probe(unsigned long long bytes, double *ms);

/* Self-check of the device square root that model code's sqrt/sqrtf use:
 * over the float bit patterns [first, first + count) (at most the 2^32 of
 * them), mismatches[0] counts inputs where it differs bitwise from the
 * compiler's IEEE float square root, mismatches[1] from the double square
 * root rounded to float; first_bad[k] is the smallest such input
 * (0xFFFFFFFF: none).  Synchronises.  No reference counterpart. */
int tay_gpu_sqrt_selftest(unsigned long long first, unsigned long long count, unsigned long long *mismatches,
                          unsigned *first_bad);

/* ------------------------------------------------------------ programs */

typedef struct TayGpuPassDesc {
    int type;                   /* TayGpuPassType */
    const char *func_name;      /* model function to call */
    TayGpuSpace *seer;          /* the act group for an act pass */
    TayGpuSpace *seen;          /* see passes only */
    float radii[4];             /* see passes only */
    int self_see;               /* see passes only; same-space pairs (i, i) skipped when 0 */
    unsigned context_size;      /* bytes copied by tay_gpu_program_set_context */
    int see_order;              /* see passes over a Z-grid: 0 the reference sweep order, 1 the balanced
                                   order (TaySeeOrder, tay_ext.h) */
    unsigned split_lanes;       /* see passes: 1, or 2, 4 or 8 lanes per seer (tay_set_see_split) */
    const char *split_clear;    /* with split_lanes > 1: the model's clear and combine functions */
    const char *split_combine;
} TayGpuPassDesc;

/* Composes the model sources and one generated kernel per pass into a single
 * hipRTC program for gfx950 and loads it.  On a compile error the hipRTC log
 * goes to log (when given) and NULL is returned.  count_pairs makes every see
 * kernel also count the pairs that pass the box test.  Replaces
 * ocl_compile_program + ocl_get_pass_kernel (tay/ocl.c:100-177, 28-33). */
TayGpuProgram *tay_gpu_program_create(const char *const *sources, unsigned sources_count,
                                      const TayGpuPassDesc *passes, unsigned passes_count,
                                      int count_pairs, char *log, unsigned log_size);
void tay_gpu_program_destroy(TayGpuProgram *program);

/* Shape of a pass without device storage, for compiling a program where no
 * GPU is present. */
typedef struct TayGpuPassShape {
    int type;                   /* TayGpuPassType */
    const char *func_name;
    unsigned seer_agent_size;   /* the act group's agent size for an act pass */
    unsigned seen_agent_size;
    int seer_is_point;
    int seen_is_point;
    int seen_kind;              /* TayGpuSpaceKind */
    int dims;
    int skip_self;              /* 1 when seer and seen are one group and self_see is 0 */
    unsigned context_size;      /* bytes of the pass context */
    int see_order;              /* as TayGpuPassDesc::see_order */
    unsigned split_lanes;       /* as TayGpuPassDesc::split_lanes, split_clear, split_combine */
    const char *split_clear;
    const char *split_combine;
} TayGpuPassShape;

/* Composes and compiles a program to a gfx950 code object without loading
 * it; needs no GPU.  On a compile error the hipRTC log goes to log. */
int tay_gpu_compile_check(const char *const *sources, unsigned sources_count, const TayGpuPassShape *shapes,
                          unsigned shapes_count, char *log, unsigned log_size, unsigned long long *code_size);

/* Uploads a pass's context (context_size bytes).  Replaces
 * ocl_push_pass_contexts_non_blocking (tay/ocl.c:252-259). */
int tay_gpu_program_set_context(TayGpuProgram *program, unsigned pass_index, const void *context);

/* Queues one pass.  Replaces ocl_run_see_kernel / ocl_run_act_kernel
 * (tay/ocl.c:208-236). */
int tay_gpu_program_run_pass(TayGpuProgram *program, unsigned pass_index);

/* Seers per lane group when a see tile starts: 64 (the default), 32, 16 or
 * 8.  A wavefront always owns 64 seers; smaller groups walk tighter cell
 * boxes one after another, which pays where agents are sparse.  Results do
 * not depend on it in the reference see order; in the balanced order and
 * in split passes it is part of the order (see_order and split_lanes of
 * TayGpuPassDesc).  No reference
 * counterpart (a tuning knob of this library). */
int tay_gpu_program_set_tile_width(TayGpuProgram *program, unsigned width);

/* A lane group of at most `width` seers (0 to 64, default 64) whose cell
 * box is still too wide for one shared sweep walks per lane (every seer its
 * own cells, candidates one by one, all members at once) instead of
 * splitting into smaller groups swept one after another; 0 always splits.
 * Pays where agents are sparse.  As the tile width, part of the order only
 * in the balanced see order and in split passes.  No reference
 * counterpart. */
int tay_gpu_program_set_lane_walk(TayGpuProgram *program, unsigned width);

/* Seers per see tile (one wavefront owns a tile): 64, 32 or 16, or 0 (the
 * default) to choose by the pass's seer count: 64 while that gives every
 * SIMD two wavefronts, fewer for smaller states, whose passes are bound by
 * the sweep of their few tiles.  Part of the order only in the balanced see
 * order; a split pass always has 64 / split_lanes seers per tile.  No
 * reference counterpart. */
int tay_gpu_program_set_tile_seers(TayGpuProgram *program, unsigned seers);

/* Pairs that passed the box test (agent_see calls) since the last reset;
 * needs count_pairs.  The reference's narrow-phase telemetry counter
 * (tay/space.c:57-60). */
int tay_gpu_program_pair_count(TayGpuProgram *program, unsigned pass_index, unsigned long long *pairs);
int tay_gpu_program_reset_pair_counts(TayGpuProgram *program);

/* Device time of the pass's kernels since the last reset, in milliseconds,
 * measured with HIP events on the library stream when timing is enabled. */
int tay_gpu_program_enable_timing(TayGpuProgram *program, int enable);
int tay_gpu_program_pass_time(TayGpuProgram *program, unsigned pass_index, double *ms, unsigned *launches);

/* ------------------------------------------------------ batched ensembles
 *
 * No reference counterpart: the reference runs one state at a time
 * (README.md:13; its process-global statics, tay/space.c:266,
 * tay/ocl.c:103).  A batch holds many independent instances of one
 * single-group model (same agent type and passes; own agent count, grid,
 * radii and contexts) in one set of device columns and advances all of them
 * with one launch per pass.  Every instance computes exactly what a state of
 * its own computes with TAY_GPU_Z_GRID and the same cell sizes: the same
 * sort, the same pairs in the same order, bit for bit.
 */

typedef struct TayGpuBatch TayGpuBatch;

typedef struct TayGpuBatchPass {
    int type;                   /* TayGpuPassType */
    const char *func_name;
    int self_see;               /* see passes: pairs (i, i) included when 1 */
    unsigned context_size;      /* bytes of one instance's context */
} TayGpuBatchPass;

/* counts[b] agents of agent_size bytes per instance; cell_table_bytes[b]
 * bounds instance b's cell table (8 bytes per cell) as shared_bytes does for
 * a space.  Compiles the model sources (hipRTC) with one batched kernel per
 * pass.  Part sizes default to 1 and radii to 0 until set. */
TayGpuBatch *tay_gpu_batch_create(const char *const *sources, unsigned sources_count,
                                  const TayGpuBatchPass *passes, unsigned passes_count, unsigned agent_size,
                                  int dims, unsigned instances, const unsigned *counts,
                                  const unsigned long long *cell_table_bytes, int count_pairs, char *log,
                                  unsigned log_size);
void tay_gpu_batch_destroy(TayGpuBatch *batch);

/* Composes and compiles a batch's program (with pair counting) without a
 * GPU. */
int tay_gpu_batch_compile_check(const char *const *sources, unsigned sources_count, const TayGpuBatchPass *passes,
                                unsigned passes_count, unsigned agent_size, int dims, char *log, unsigned log_size,
                                unsigned long long *code_size);

/* instances x 4 floats: suggested cell sizes (min_part_sizes) per instance */
int tay_gpu_batch_set_part_sizes(TayGpuBatch *batch, const float *part_sizes);
/* instances x 4 floats: the part sizes in use (1 until set) */
int tay_gpu_batch_get_part_sizes(const TayGpuBatch *batch, float *part_sizes);
/* instances x 4 floats: part sizes for about `target` agents per cell.  An
 * instance of n agents whose current bounding box has volume V (area or
 * length in fewer dimensions) gets cells of side (target V / n)^(1/dims),
 * clamped per axis to [min_factor, max_factor] times its radius of the
 * first see pass.  Computes them without applying them. */
int tay_gpu_batch_density_part_sizes(TayGpuBatch *batch, float target, float min_factor, float max_factor,
                                     float *part_sizes);
/* instances x 4 floats: the see pass's radii per instance */
int tay_gpu_batch_set_radii(TayGpuBatch *batch, unsigned pass_index, const float *radii);
/* instances x context_size bytes, packed: the pass's context per instance */
int tay_gpu_batch_set_contexts(TayGpuBatch *batch, unsigned pass_index, const void *contexts);
/* seers per lane group at the start of a see tile: 64, 32, 16 or 8 */
int tay_gpu_batch_set_tile_width(TayGpuBatch *batch, unsigned width);

/* as tay_gpu_program_set_lane_walk */
int tay_gpu_batch_set_lane_walk(TayGpuBatch *batch, unsigned width);

/* The order in which the persistent see kernels take tiles: estimated cost,
 * heaviest first (the default), or storage order (neighbouring tiles share
 * seen cells in cache).  Results do not depend on it. */
typedef enum TayGpuTileOrder {
    TAY_GPU_TILE_ORDER_COST = 0,
    TAY_GPU_TILE_ORDER_STORAGE = 1,
} TayGpuTileOrder;
int tay_gpu_batch_set_tile_order(TayGpuBatch *batch, int order);

/* Sub-cell bits of every instance's storage order, as
 * tay_gpu_space_set_order_bits. */
int tay_gpu_batch_set_order_bits(TayGpuBatch *batch, int bits);

/* All instances' agents as AoS records, instance after instance (sum of
 * counts records). */
int tay_gpu_batch_push(TayGpuBatch *batch, const void *agents);
int tay_gpu_batch_fetch(TayGpuBatch *batch, void *agents);

/* Runs `steps` steps (sort every instance, then the passes in order).
 * use_graph captures one step into a HIP graph (once, and again after a
 * change of tile width) and replays it; otherwise kernels are launched one
 * by one, and pass timing, when enabled, records each launch.  No host
 * synchronisation; *device_ms (optional) receives the device time of the
 * whole run, which synchronises. */
int tay_gpu_batch_run(TayGpuBatch *batch, unsigned steps, int use_graph, double *device_ms);

int tay_gpu_batch_enable_timing(TayGpuBatch *batch, int enable);
int tay_gpu_batch_pass_time(TayGpuBatch *batch, unsigned pass_index, double *ms, unsigned *launches);
int tay_gpu_batch_pair_count(TayGpuBatch *batch, unsigned pass_index, unsigned long long *pairs);
int tay_gpu_batch_reset_pair_counts(TayGpuBatch *batch);

/* Autotuner: for every cell factor f (cell size = f x the instance's radius
 * of the first see pass, per axis), tile width and lane-walk width, restores
 * the agents it started from, runs `steps` un-captured steps and times them
 * on the device.  table_ms (optional, factors x widths x lane walks, lane
 * walk fastest) receives ms per step.  Applies the fastest combination,
 * reports it and leaves the agents as they were.  The cell size changes the
 * grid, and so the order of each seer's pairs; the widths do not. */
int tay_gpu_batch_autotune(TayGpuBatch *batch, const float *cell_factors, unsigned factors_count,
                           const unsigned *widths, unsigned widths_count, const unsigned *lane_walks,
                           unsigned lane_walks_count, unsigned steps, double *table_ms, float *best_factor,
                           unsigned *best_width, unsigned *best_lane_walk);

/* As tay_gpu_batch_autotune, over agents-per-cell targets instead of one
 * cell factor for every instance: each target gives every instance its own
 * cell size by its density (tay_gpu_batch_density_part_sizes, densities of
 * the starting agents), so dense and sparse instances of one ensemble get
 * fine and coarse grids.  Applies the fastest combination and writes the
 * chosen target and (optional, instances x 4 floats) its part sizes. */
int tay_gpu_batch_autotune_density(TayGpuBatch *batch, const float *targets, unsigned targets_count,
                                   float min_factor, float max_factor, const unsigned *widths, unsigned widths_count,
                                   const unsigned *lane_walks, unsigned lane_walks_count, unsigned steps,
                                   double *table_ms, float *best_target, float *part_sizes, unsigned *best_width,
                                   unsigned *best_lane_walk);

#ifdef __cplusplus
}
#endif

#endif
