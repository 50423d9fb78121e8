/*
 * device_abi.h - structures shared by the ahead-of-time grid kernels
 * (grid_kernels.hip), the host code and the kernels generated at run time.
 * The build embeds this file's text at the top of every hipRTC program, so it
 * must compile both under hipcc and under hipRTC (no includes).
 */
#ifndef TAY_DEVICE_ABI_H
#define TAY_DEVICE_ABI_H

/* Grid geometry written by the geometry kernel at every sort and read by the
 * key, range, permute and see kernels.  Unused axes have count 1. */
struct TayGridGeom {
    float origin[4];
    float cell_size[4];
    int count[4];
    unsigned long long cells;   /* product of count over the used axes */
    float box_min[4];
    float box_max[4];
};

/* One cell's agents are [beg, end) of the key-sorted storage. */
struct TayCellRange {
    unsigned beg;
    unsigned end;
};

/* Arguments of a generated see kernel.  Agent storage is structure-of-arrays:
 * word w of agent i is cols[w * cap + i]. */
struct TaySeeArgs {
    const unsigned *seer_cols;
    unsigned *seer_out;             /* same storage as seer_cols: seers update their own words */
    unsigned seer_cap;
    unsigned seer_count;
    const unsigned *seen_cols;
    unsigned seen_cap;
    unsigned seen_count;
    const struct TayCellRange *ranges;  /* seen grid; unused for a Simple seen space */
    const struct TayGridGeom *geom;
    const void *context;
    float radii[4];
    unsigned long long *pair_count;     /* pairs passing the box test, when counting */
    unsigned tiles;                     /* number of 64-seer tiles */
};

struct TayActArgs {
    unsigned *cols;
    unsigned cap;
    unsigned count;
    const void *context;
};

#endif
