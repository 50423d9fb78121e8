/*
 * oracle.h - TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's Tay library for the Z-grid hot path.
 * It is the checker for the GPU library and the timed CPU baseline; nothing in
 * the product (tay_amd/, include/) may load it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * liboracle.so exports the reference API of include/tay.h (same names and
 * signatures) with these spaces:
 *   TAY_CPU_SIMPLE  all pairs                (tay/cpu_simple.c)
 *   TAY_CPU_GRID    uniform grid             (tay/cpu_grid.c)
 *   TAY_CPU_Z_GRID  power-of-two Morton grid (tay/cpu_z_grid.c)
 *   TAY_GPU_Z_GRID  the CPU statement of what the GPU Z-grid must compute
 *                   (geometry of the reference's OpenCL Z-grid,
 *                   tay/ocl_grid.c:128-136, Morton keys of tay/cpu_z_grid.c:37-83,
 *                   stable key order, integer memory cap, x->y->z sweep)
 * plus the inspection calls below, which have no reference counterpart.
 */
#ifndef tay_oracle_h
#define tay_oracle_h

#include "tay.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Pairs of the pass summed over the steps of the last tay_run: candidates
 * tested (broad) and pairs that passed the box test, i.e. agent_see calls
 * (narrow); the reference's telemetry counters (tay/space.c:50-60,119-121). */
void oracle_get_pair_counts(TayState *state, unsigned pass_index,
                            unsigned long long *broad, unsigned long long *narrow);

/* Geometry of a grid space after its last sort. */
void oracle_get_grid_info(TayState *state, TayGroup *group,
                          float4 *origin, float4 *cell_sizes, int *cell_counts /* [4] */,
                          unsigned long long *cells_count);

/* Morton key of cell indices (tay/cpu_z_grid.c:37-83): 3D 10 bits per axis,
 * 2D 16 bits per axis, 1D identity, 4D not implemented (returns 0). */
unsigned oracle_morton_key(int x, int y, int z, int dims);
void oracle_morton_decode(unsigned key, int dims, int *x, int *y, int *z);

/* Sub-cell bits (0 to 3, default 2) of the GPU Z-grid statement's storage
 * order inside a cell; the GPU library exports the same call. */
int tay_set_order_bits(TayState *state, TayGroup *group, int bits);

/* See order of the GPU Z-grid statement (0: the reference sweep order, the
 * default; 1: the balanced order of the GPU library's see kernel, whose
 * tile layout the next three calls set as the GPU library's calls of the
 * same names do: seers per tile (0: by the seer count, for a device of 1024
 * SIMDs), starting lane-group width, lane-walk width).  The GPU library
 * exports the same calls (include/tay_ext.h). */
int tay_set_see_order(TayState *state, int order);
/* A split see pass (as the GPU library's call, include/tay_ext.h): every
 * seer's pairs go to `lanes` (1, 2, 4 or 8) partial records by their
 * position in the lane group's sweep, merged at the end of the tile; the
 * oracle calls the CPU callbacks clear and combine. */
int tay_set_see_split(TayState *state, unsigned pass_index, unsigned lanes, void (*clear)(void *),
                      const char *clear_name, void (*combine)(void *, void *), const char *combine_name);
int tay_set_tile_seers(TayState *state, unsigned seers);
int tay_set_tile_width(TayState *state, unsigned width);
int tay_set_lane_walk(TayState *state, unsigned width);

/* Wall-clock seconds spent in see passes during the last tay_run. */
double oracle_get_see_seconds(TayState *state);

/* A see pass that runs only its first `count` seers (in storage order after
 * the sort; 0: all), so a single-thread baseline can time a bounded sample of
 * a large step: pairs and see seconds then cover that sample. */
void oracle_limit_seers(TayState *state, unsigned pass_index, unsigned count);

#ifdef __cplusplus
}
#endif

#endif
