/*
 * device_abi.h - structures shared by the ahead-of-time grid kernels
 * (grid_build.h, used by gpu_space.hip, gpu_batch.hip and dflock.hip), the
 * host code and the kernels generated at run time.
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
    int order_bits;             /* sub-cell bits per axis of the storage-order key */
    /* largest extent (max corner - min corner) of the group's agents per
     * axis; 0 for point agents.  Non-point agents are binned by their min
     * corner, so a search for agents whose box reaches a region starts this
     * far below it. */
    float max_extent[4];
};

/* A see pass counts its pairs into this many 64-bit counters, wavefronts of
 * block b adding to counter b mod TAY_PAIR_SLOTS: one counter served every
 * wavefront's atomic in turn, which took longer than a short pass itself
 * (125,000 wavefronts of a 2M-particle density pass split over 4 lanes:
 * 1.55 ms counting against 0.56 ms not). */
#define TAY_PAIR_SLOTS 64u

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
    unsigned long long *pair_count;     /* pairs passing the box test, when counting: TAY_PAIR_SLOTS
                                           counters, summed when read */
    unsigned tiles;                     /* number of tiles (tile_seers seers each) */
    unsigned width;                     /* seers per lane group at the start of a tile: 64, 32, 16 or 8 */
    unsigned lane_walk;                 /* sparse groups of at most this many lanes walk per lane */
    unsigned tile_seers;                /* seers per tile (one wavefront): 64, 32 or 16 */
};

struct TayActArgs {
    unsigned *cols;
    unsigned cap;
    unsigned count;
    const void *context;
};

/* One instance of a batched ensemble.  Its agents occupy slots
 * [off, off + count) of the shared columns; off is a multiple of 64, so
 * every 64-slot tile belongs to one instance, and the slots up to the next
 * multiple of 64 are padding. */
struct TayBatchInstance {
    unsigned off;
    unsigned count;
    unsigned rec;                       /* first record of the instance in host (AoS) order */
    unsigned pad;
    unsigned long long ranges_off;      /* first entry of its cell table */
    unsigned long long max_cells;       /* entries of its cell table */
};

/* Arguments of a generated batched see kernel: one persistent launch for
 * every instance.  Wavefronts take tiles from a work queue in the order given
 * by `order` (heaviest first). */
struct TayBatchSeeArgs {
    unsigned *cols;
    unsigned cap;
    unsigned tiles;
    const unsigned *tile_inst;          /* instance of each tile */
    const unsigned *order;              /* tiles in processing order */
    unsigned *queue;                    /* work-queue head, zero at launch */
    const struct TayBatchInstance *inst;
    const struct TayGridGeom *geom;     /* one per instance */
    const struct TayCellRange *ranges;  /* all cell tables */
    const float *radii;                 /* 4 per instance */
    const unsigned *context;            /* ctx_words per instance */
    unsigned ctx_words;
    unsigned width;
    unsigned lane_walk;
    unsigned long long *pair_count;
};

struct TayBatchActArgs {
    unsigned *cols;
    unsigned cap;
    unsigned slots;
    const unsigned *tile_inst;
    const struct TayBatchInstance *inst;
    const unsigned *context;
    unsigned ctx_words;
};

#endif
