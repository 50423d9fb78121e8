/*
 * tay_ext.h - additions to the Tay API in libtay_amd.so.  None of these has a
 * counterpart in the reference; they expose the GPU space's telemetry and let
 * hosts add many agents at once.
 */
#ifndef tay_ext_h
#define tay_ext_h

#include "tay.h"
#include "tay_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last error set on the state ("" if none). */
const char *tay_get_error_message(TayState *state);

/* Commits the next count agents of the group's storage at once (the
 * records at tay_get_available_agent() onwards); same as calling
 * tay_commit_available_agent count times. */
void tay_commit_available_agents(TayState *state, TayGroup *group, unsigned count);

/* Counting of pairs that pass the box test (agent_see calls) in every see
 * pass; takes effect at the next tay_run, which recompiles the program.  The
 * counts cover the steps of the last tay_run. */
void tay_set_pair_counting(TayState *state, int enable);
int tay_get_pair_count(TayState *state, unsigned pass_index, unsigned long long *pairs);

/* Device time of each pass's kernels over the last tay_run, measured with
 * HIP events on the library stream. */
void tay_set_pass_timing(TayState *state, int enable);
int tay_get_pass_time(TayState *state, unsigned pass_index, double *ms, unsigned *launches);

/* Order of a Z-grid group's agents inside each cell: refined by `bits`
 * sub-cell bits per axis (0 to 3; default 2), i.e. by the Morton order of a
 * 2^bits-finer grid, then by previous storage order; 0 keeps the previous
 * storage order alone, as the reference's CPU grids do.  The finest order
 * whose key fits the Morton range is used each step.  It changes the order
 * in which a seer meets its pairs (not the pairs), so float sums differ in
 * rounding between settings. */
int tay_set_order_bits(TayState *state, TayGroup *group, int bits);

/* Seers per lane group when a see tile starts (64, 32, 16 or 8; default
 * 64), as tay_gpu_program_set_tile_width.  In the reference see order
 * results do not depend on it; in the balanced order and in split passes
 * (tay_set_see_split) it is part of the order in which each seer meets its
 * pairs. */
int tay_set_tile_width(TayState *state, unsigned width);

/* Lane-walk width of the see kernels (0 to 64), as
 * tay_gpu_program_set_lane_walk; as the tile width, part of the order in
 * the balanced order and in split passes only. */
int tay_set_lane_walk(TayState *state, unsigned width);

/* Seers per see tile (0 = by the seer count, the default; 64, 32 or 16),
 * as tay_gpu_program_set_tile_seers; part of the order in the balanced
 * order only (a split pass has 64 / lanes seers per tile whatever this
 * says). */
int tay_set_tile_seers(TayState *state, unsigned seers);

/* Order in which a seer meets its pairs in see passes over a GPU Z-grid.
 * TAY_SEE_ORDER_REFERENCE (the default) is the reference CPU grids' sweep
 * (cells x, then y, then z, each in storage order; tay/cpu_grid.c:222-261),
 * so float sums are bit-identical to a CPU grid's with the same storage
 * order.  TAY_SEE_ORDER_BALANCED takes the candidates of a tile's cells in
 * strided chunks, each an even sample of the tile's neighbourhood, which
 * keeps more lanes busy in the see kernel: the same pairs and neighbour
 * counts, float sums in another fixed order, set by the seer count, the
 * tile seers, the tile width and the lane walk.  Takes effect at the next
 * tay_run (the program is compiled again).  Returns 0 for an unknown
 * order.  No reference counterpart. */
typedef enum TaySeeOrder {
    TAY_SEE_ORDER_REFERENCE = 0,
    TAY_SEE_ORDER_BALANCED = 1,
} TaySeeOrder;
int tay_set_see_order(TayState *state, int order);

/* Splits each seer of see pass `pass_index` (its place among the state's
 * passes) over `lanes` (1, 2, 4 or 8; 1 undoes the split) lanes of the see
 * kernel, for a model whose see function only adds up its pairs'
 * contributions.  Lane k of a seer takes the candidates at positions
 * p = k (mod lanes) of its lane group's sweep (cells x, then y, then z,
 * each in storage order), in that order, into a partial record: partial 0
 * is the seer itself, partial k > 0 a copy emptied by `clear`; at the end
 * of the tile the partials merge by `combine(a, b)` (a += b) in a fixed
 * tree: pairs (0 + 1), (2 + 3), ..., then pairs of those, ((0 + 1) + (2 +
 * 3)) for 4 lanes.  A wavefront then holds 64 / lanes seers and its compact
 * rounds run each seer's pairs on `lanes` lanes: the same pairs and counts,
 * float sums in that fixed order.  A lane group whose cell box is too wide
 * halves down to single seers, whatever tay_set_lane_walk says (a lane walk
 * would use one lane of each seer).  The function pointers are the
 * CPU callbacks (tay_add_see's convention; unused by this library), the
 * names the model functions to compile.  Such a pass runs in the reference
 * see order whatever tay_set_see_order says.  Takes effect at the next
 * tay_run.  No reference counterpart. */
int tay_set_see_split(TayState *state, unsigned pass_index, unsigned lanes, void (*clear)(void *),
                      const char *clear_name, void (*combine)(void *, void *), const char *combine_name);

/* The device space of a group, valid while the simulation runs (NULL before
 * the first tay_run and for groups not on the GPU). */
TayGpuSpace *tay_get_gpu_space(TayState *state, TayGroup *group);

/* When agents come back to host storage.  TAY_FETCH_EVERY_RUN (default) is
 * the reference's behaviour: every tay_run ends by copying all agents of all
 * groups to the host (tay/state.c:313-315).  TAY_FETCH_ON_DEMAND leaves them
 * on the device until tay_fetch_agents() is called, so short runs in an
 * interactive loop do not pay a full copy each time. */
typedef enum TayFetchMode {
    TAY_FETCH_EVERY_RUN = 0,
    TAY_FETCH_ON_DEMAND = 1,
} TayFetchMode;

void tay_set_fetch_mode(TayState *state, int mode);

/* Copies the group's agents to host storage now (all groups when group is
 * NULL); returns 1 on success.  Before the first tay_run and after
 * tay_simulation_end the host storage already holds the newest agents
 * (the end of a simulation fetches them under TAY_FETCH_ON_DEMAND). */
int tay_fetch_agents(TayState *state, TayGroup *group);

/* Copies bytes [offset, offset + size) of each of the group's agents, in
 * current storage order, into dst as packed records of `size` bytes (for
 * example offset 8, size 12: the positions), leaving host storage alone.
 * offset and size are multiples of 4 within the agent size; the group must
 * be on the device (after a tay_run).  Returns the number of records
 * copied, 0 on error. */
unsigned tay_fetch_agent_fields(TayState *state, TayGroup *group, unsigned offset, unsigned size, void *dst);

#ifdef __cplusplus
}
#endif

#endif
