/*
 * gpu_batch.hip - batched ensembles: many independent instances of one
 * single-group model advanced together, one launch per pass.
 *
 * Layout.  All instances share one set of structure-of-arrays columns.
 * Instance b owns slots [off_b, off_b + count_b) with off_b a multiple of 64,
 * so every 64-slot tile belongs to one instance (tile_inst[tile]); the slots
 * up to the next multiple of 64 are padding that no kernel reads as agents.
 * Each instance has its own grid geometry and its own cell table
 * (ranges + ranges_off_b).
 *
 * Sort, per step, for every instance at once:
 *   box       one wavefront per tile: min/max of its positions
 *   geometry  one wavefront per instance: reduce its tiles' boxes, then the
 *             single-space geometry; the largest key any grid allows
 *   keys      per slot: the instance-local Morton key, padding 0xFFFFFFFF
 *   radix     stable passes over the key bytes on radix tiles that never
 *             cross an instance, histograms laid out instance-major, so one
 *             scan places every key inside its instance: key order, and
 *             within a key the previous storage order -- exactly the order
 *             a sort of the instance alone produces; a pass over digits
 *             above the largest key the grids allow only copies
 *             (radix_pass_idle)
 *   ranges    clear and fill each instance's table (absolute slots)
 *   permute   gather the columns, write the tag
 *   order     tiles bucketed by estimated cost (occupancy of the cell of the
 *             tile's middle seer x its seers), heaviest bucket first; the
 *             persistent see kernels take tiles from a work queue in this
 *             order, so the longest tiles start first
 * No kernel of a step needs host synchronisation or a memset, so a step is
 * captured into a HIP graph once and replayed.
 */
#include "internal.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

using namespace tay;

#include "grid_build.h"

namespace {

constexpr unsigned kCostBuckets = 32;

/* One wavefront per tile: bounding box of the tile's agents. */
__global__ __launch_bounds__(kBuildBlock) void k_tile_box(const unsigned *__restrict__ cols, unsigned cap,
                                                          unsigned tiles, const unsigned *__restrict__ tile_inst,
                                                          const TayBatchInstance *__restrict__ inst, int dims,
                                                          float *__restrict__ partials,
                                                          unsigned *__restrict__ key_max) {
    if (blockIdx.x == 0 && threadIdx.x == 0)
        *key_max = 0u;      /* raised by k_batch_geometry */
    const unsigned t = blockIdx.x * (kBuildBlock / 64u) + (threadIdx.x >> 6);
    if (t >= tiles)
        return;
    const unsigned lane = threadIdx.x & 63u;
    const unsigned i = t * 64u + lane;
    const TayBatchInstance in = inst[tile_inst[t]];
    const bool real = i - in.off < in.count;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float p = real && k < dims ? __uint_as_float(cols[(2u + k) * cap + i]) : 0.0f;
        const float lo = wave_min(real ? p : FLT_MAX), hi = wave_max(real ? p : -FLT_MAX);
        if (lane == 0) {
            partials[t * 6u + k] = lo;
            partials[t * 6u + 3 + k] = hi;
        }
    }
}

/* One wavefront per instance: reduce its tiles' boxes, then lane 0 derives
 * the grid as tay_gpu_space_sort does, and raises key_max (zeroed by
 * k_tile_box) to the largest key the grid allows: the Morton key of the
 * last sub-cell on every axis, at least any agent's key. */
__global__ __launch_bounds__(64) void k_batch_geometry(const float *__restrict__ partials,
                                                       const float *__restrict__ parts,
                                                       const TayBatchInstance *__restrict__ inst, int dims,
                                                       int order_bits, TayGridGeom *__restrict__ geom,
                                                       unsigned *__restrict__ key_max) {
    const unsigned b = blockIdx.x, lane = threadIdx.x;
    const TayBatchInstance in = inst[b];
    const unsigned t0 = in.off / 64u, t1 = (in.off + in.count + 63u) / 64u;
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (unsigned t = t0 + lane; t < t1; t += 64u) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(lo[k], partials[t * 6u + k]);
            hi[k] = fmaxf(hi[k], partials[t * 6u + 3 + k]);
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = wave_min(lo[k]);
        hi[k] = wave_max(hi[k]);
    }
    if (lane != 0)
        return;
    float part[3];
    for (int k = 0; k < 3; ++k) {
        part[k] = parts[b * 4u + k];
        if (in.count == 0)
            lo[k] = hi[k] = 0.0f;   /* no agents: a one-cell grid that no tile reads */
    }
    grid_geometry(lo, hi, part, dims, in.max_cells, order_bits, geom + b);
    const TayGridGeom &g = geom[b];
    int last[3] = {0, 0, 0};
    for (int k = 0; k < dims; ++k)
        last[k] = ((g.count[k] - 1) << g.order_bits) | ((1 << g.order_bits) - 1);
    atomicMax(key_max, morton(last[0], last[1], last[2], dims));
}

__global__ __launch_bounds__(kBuildBlock) void k_batch_keys(const unsigned *__restrict__ cols, unsigned cap,
                                                            unsigned slots, const unsigned *__restrict__ tile_inst,
                                                            const TayBatchInstance *__restrict__ inst, int dims,
                                                            const TayGridGeom *__restrict__ geom,
                                                            unsigned *__restrict__ keys) {
    const unsigned i = blockIdx.x * kBuildBlock + threadIdx.x;
    if (i >= slots)
        return;
    const unsigned b = tile_inst[i >> 6];
    if (i - inst[b].off >= inst[b].count) {
        keys[i] = 0xFFFFFFFFu;   /* padding sorts after every cell of its instance */
        return;
    }
    keys[i] = storage_key(cols, cap, i, dims, geom + b);
}

/* blockIdx.y is the instance; its blocks clear its table's used cells. */
__global__ __launch_bounds__(kBuildBlock) void k_batch_ranges_clear(TayCellRange *__restrict__ ranges,
                                                                    const TayBatchInstance *__restrict__ inst,
                                                                    const TayGridGeom *__restrict__ geom) {
    const unsigned b = blockIdx.y;
    TayCellRange *r = ranges + inst[b].ranges_off;
    const unsigned long long cells = geom[b].cells;
    for (unsigned long long c = blockIdx.x * (unsigned long long)kBuildBlock + threadIdx.x; c < cells;
         c += (unsigned long long)gridDim.x * kBuildBlock) {
        r[c].beg = 0;
        r[c].end = 0;
    }
}

__global__ __launch_bounds__(kBuildBlock) void k_batch_ranges(const unsigned *__restrict__ keys, unsigned slots,
                                                              const unsigned *__restrict__ tile_inst,
                                                              const TayBatchInstance *__restrict__ inst, int dims,
                                                              const TayGridGeom *__restrict__ geom,
                                                              TayCellRange *__restrict__ ranges) {
    const unsigned i = blockIdx.x * kBuildBlock + threadIdx.x;
    if (i >= slots)
        return;
    const unsigned b = tile_inst[i >> 6];
    const TayBatchInstance in = inst[b];
    const unsigned li = i - in.off;
    if (li >= in.count)
        return;
    const TayGridGeom *g = geom + b;
    const unsigned key = cell_of_key(keys[i], dims, g);
    const bool first = li == 0 || cell_of_key(keys[i - 1], dims, g) != key;
    const bool last = li + 1 == in.count || cell_of_key(keys[i + 1], dims, g) != key;
    if (first || last) {
        TayCellRange *r = ranges + in.ranges_off + morton_to_cell(key, dims, g);
        if (first)
            r->beg = i;
        if (last)
            r->end = i + 1;
    }
}

/* Gather every column into sorted order; the tag gets the key and the rank
 * inside the cell, as k_permute writes it. */
__global__ __launch_bounds__(kBuildBlock) void k_batch_permute(const unsigned *__restrict__ keys,
                                                               const unsigned *__restrict__ vals, unsigned slots,
                                                               unsigned words, unsigned cap,
                                                               const unsigned *__restrict__ tile_inst,
                                                               const TayBatchInstance *__restrict__ inst, int dims,
                                                               const TayGridGeom *__restrict__ geom,
                                                               const TayCellRange *__restrict__ ranges,
                                                               const unsigned *__restrict__ src,
                                                               unsigned *__restrict__ dst) {
    const unsigned i = blockIdx.x * kBuildBlock + threadIdx.x;
    if (i >= slots)
        return;
    const unsigned b = tile_inst[i >> 6];
    const TayBatchInstance in = inst[b];
    const unsigned key = keys[i];
    const unsigned from = vals[i];
    const bool real = i - in.off < in.count;
    dst[i] = real ? cell_of_key(key, dims, geom + b) : key;
    dst[cap + i] = real ? i - ranges[in.ranges_off + key_to_cell(key, dims, geom + b)].beg : 0u;
    gather_words(src, dst, cap, from, i, 2u, words);
}

/* Estimated cost of each tile, bucketed by its binary logarithm; the block
 * counts its tiles per bucket in LDS and adds the counts to the histogram. */
__global__ __launch_bounds__(kBuildBlock) void k_tile_cost(const unsigned *__restrict__ keys, unsigned tiles,
                                                           const unsigned *__restrict__ tile_inst,
                                                           const TayBatchInstance *__restrict__ inst, int dims,
                                                           const TayGridGeom *__restrict__ geom,
                                                           const TayCellRange *__restrict__ ranges,
                                                           const float *__restrict__ radii,
                                                           unsigned *__restrict__ tile_bucket,
                                                           unsigned *__restrict__ bucket_hist) {
    __shared__ unsigned h[kCostBuckets];
    if (threadIdx.x < kCostBuckets)
        h[threadIdx.x] = 0u;
    __syncthreads();
    const unsigned t = blockIdx.x * kBuildBlock + threadIdx.x;
    if (t < tiles) {
        const unsigned b = tile_inst[t];
        const TayBatchInstance in = inst[b];
        const unsigned s0 = t * 64u;
        const unsigned n = in.off + in.count - s0 < 64u ? in.off + in.count - s0 : 64u;
        const unsigned mid = s0 + n / 2u;
        const TayCellRange r = ranges[in.ranges_off + key_to_cell(keys[mid], dims, geom + b)];
        /* candidates per seer: the cell's agents times the cells a seer's
         * box spans, so instances with different cell sizes compare */
        float span = 1.0f;
        if (radii)
            for (int k = 0; k < dims; ++k)
                span *= 2.0f * radii[4u * b + (unsigned)k] / geom[b].cell_size[k] + 1.0f;
        const float c = (float)(r.end - r.beg) * (float)n * span;
        const unsigned cost = c < 4.0e9f ? (unsigned)c : 0xFFFFFFFFu;
        const unsigned bucket = cost ? 31u - (unsigned)__clz(cost) : 0u;
        tile_bucket[t] = bucket;
        atomicAdd(&h[bucket], 1u);
    }
    __syncthreads();
    if (threadIdx.x < kCostBuckets && h[threadIdx.x])
        atomicAdd(&bucket_hist[threadIdx.x], h[threadIdx.x]);
}

/* One wavefront: bucket offsets, heaviest bucket first; clears the histogram
 * for the next step and the work queues of this step's see passes. */
__global__ __launch_bounds__(64) void k_tile_scan(unsigned *__restrict__ bucket_hist,
                                                  unsigned *__restrict__ bucket_off, unsigned *__restrict__ queues,
                                                  unsigned queues_count) {
    const unsigned lane = threadIdx.x;
    /* lane l holds bucket 31 - l, so an inclusive scan runs heaviest first */
    const unsigned v = lane < kCostBuckets ? bucket_hist[kCostBuckets - 1u - lane] : 0u;
    const unsigned inc = wave_inclusive_sum(v, lane);
    if (lane < kCostBuckets) {
        bucket_off[kCostBuckets - 1u - lane] = inc - v;
        bucket_hist[kCostBuckets - 1u - lane] = 0u;
    }
    if (lane < queues_count)
        queues[lane] = 0u;
}

/* Places each tile in its bucket's run of the order: ranks inside the block
 * from LDS counters, one reservation per bucket and block.  The order inside
 * a bucket is not fixed, which no result depends on. */
__global__ __launch_bounds__(kBuildBlock) void k_tile_scatter(const unsigned *__restrict__ tile_bucket,
                                                              unsigned tiles, unsigned *__restrict__ bucket_off,
                                                              unsigned *__restrict__ order) {
    __shared__ unsigned count[kCostBuckets];
    __shared__ unsigned base[kCostBuckets];
    if (threadIdx.x < kCostBuckets)
        count[threadIdx.x] = 0u;
    __syncthreads();
    const unsigned t = blockIdx.x * kBuildBlock + threadIdx.x;
    const unsigned bucket = t < tiles ? tile_bucket[t] : 0u;
    const unsigned rank = t < tiles ? atomicAdd(&count[bucket], 1u) : 0u;
    __syncthreads();
    if (threadIdx.x < kCostBuckets)
        base[threadIdx.x] = count[threadIdx.x] ? atomicAdd(&bucket_off[threadIdx.x], count[threadIdx.x]) : 0u;
    __syncthreads();
    if (t < tiles)
        order[base[bucket] + rank] = t;
}

/* Storage order: tile k is processed k-th; also clears the work queues. */
__global__ __launch_bounds__(kBuildBlock) void k_tile_identity(unsigned tiles, unsigned *__restrict__ order,
                                                               unsigned *__restrict__ queues, unsigned queues_count) {
    const unsigned t = blockIdx.x * kBuildBlock + threadIdx.x;
    if (t < tiles)
        order[t] = t;
    if (t < queues_count)
        queues[t] = 0u;
}

/* Instance holding record r of the host order (binary search over rec). */
__device__ __forceinline__ unsigned instance_of_record(const TayBatchInstance *__restrict__ inst, unsigned instances,
                                                       unsigned r) {
    unsigned lo = 0, hi = instances;
    while (hi - lo > 1u) {
        const unsigned mid = (lo + hi) / 2u;
        if (inst[mid].rec <= r)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

/* AoS records (instance after instance) <-> the padded columns. */
__global__ __launch_bounds__(kBuildBlock) void k_batch_transpose(unsigned *__restrict__ aos,
                                                                 unsigned *__restrict__ cols, unsigned records,
                                                                 unsigned words, unsigned cap,
                                                                 const TayBatchInstance *__restrict__ inst,
                                                                 unsigned instances, int to_cols) {
    const unsigned long long total = (unsigned long long)records * words;
    for (unsigned long long e = blockIdx.x * (unsigned long long)kBuildBlock + threadIdx.x; e < total;
         e += (unsigned long long)gridDim.x * kBuildBlock) {
        const unsigned r = (unsigned)(e / words), w = (unsigned)(e % words);
        const unsigned b = instance_of_record(inst, instances, r);
        const unsigned slot = inst[b].off + (r - inst[b].rec);
        if (to_cols)
            cols[(unsigned long long)w * cap + slot] = aos[e];
        else
            aos[e] = cols[(unsigned long long)w * cap + slot];
    }
}

}  // namespace

/* ---------------------------------------------------------------- state */

struct TayGpuBatch {
    struct Pass {
        TayGpuBatchPass desc;
        std::string func_name;
        unsigned ctx_words;
        hipFunction_t kernel;
        unsigned *context;          /* ctx_words per instance */
        float *radii;               /* 4 per instance (see passes) */
        unsigned long long *pair_count;
        unsigned grid;              /* persistent see grid */
        /* launch arguments per starting column buffer; they outlive the
         * launches and the captured graphs that refer to them */
        TayBatchSeeArgs see_args[2];
        TayBatchActArgs act_args[2];
        std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pending;
        std::vector<hipEvent_t> ev_free;
        double ms;
        unsigned launches;
    };

    unsigned agent_size, words, instances, records, slots, tiles;
    int dims;
    int count_pairs;
    int timing;
    unsigned width;
    int order_mode;             /* TayGpuTileOrder */
    int order_bits;             /* sub-cell bits of the storage order */
    unsigned lane_walk;         /* sparse lane groups up to this width walk per lane */
    unsigned sort_tiles;        /* radix tiles of the per-instance segmented sort */
    std::vector<TayBatchInstance> inst_host;
    std::vector<float> radius_host;     /* first see pass's radii, 4 per instance */
    std::vector<float> parts_host;      /* part sizes, 4 per instance */
    unsigned long long total_cells;
    unsigned long long max_instance_cells;

    unsigned *cols[2];
    int cur;
    TayBatchInstance *inst;
    unsigned *tile_inst;
    float *parts;
    float *boxes;
    TayGridGeom *geom;
    unsigned *keys[2];
    unsigned *vals[2];
    unsigned *block_hist;
    TaySortTile *sort_seg;
    unsigned *scan_sums;
    unsigned *key_max;          /* largest key the grids allow this step (the word after scan_sums) */
    TayCellRange *ranges;
    unsigned *tile_bucket;
    unsigned *bucket_hist;
    unsigned *bucket_off;
    unsigned *order;
    unsigned *queues;
    unsigned *staging;
    unsigned *backup;

    hipModule_t module;
    std::vector<Pass> passes;
    /* captured two-step graphs, one per starting column buffer */
    hipGraph_t graph[2];
    hipGraphExec_t exec[2];
};

namespace {

bool take_event(TayGpuBatch::Pass &pass, hipEvent_t *e) {
    if (!pass.ev_free.empty()) {
        *e = pass.ev_free.back();
        pass.ev_free.pop_back();
        return true;
    }
    return hip_ok(hipEventCreate(e), "hipEventCreate");
}

bool resolve_timing(TayGpuBatch::Pass &pass) {
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

void drop_graph(TayGpuBatch *b) {
    for (int k = 0; k < 2; ++k) {
        if (b->exec[k])
            (void)hipGraphExecDestroy(b->exec[k]);
        if (b->graph[k])
            (void)hipGraphDestroy(b->graph[k]);
        b->exec[k] = nullptr;
        b->graph[k] = nullptr;
    }
}

/* Queues one step.  `record` adds timing events around the passes (never
 * while capturing). */
bool batch_step(TayGpuBatch *b, hipStream_t s, bool record) {
    const unsigned n = b->slots, cap = b->slots;
    const unsigned *src = b->cols[b->cur];
    unsigned *dst = b->cols[b->cur ^ 1];
    const unsigned agent_blocks = blocks_for(n, kBuildBlock, 0xFFFFFFFFu);

    k_tile_box<<<(b->tiles + kBuildBlock / 64u - 1u) / (kBuildBlock / 64u), kBuildBlock, 0, s>>>(
        src, cap, b->tiles, b->tile_inst, b->inst, b->dims, b->boxes, b->key_max);
    k_batch_geometry<<<b->instances, 64, 0, s>>>(b->boxes, b->parts, b->inst, b->dims, b->order_bits, b->geom,
                                                 b->key_max);
    k_batch_keys<<<agent_blocks, kBuildBlock, 0, s>>>(src, cap, n, b->tile_inst, b->inst, b->dims, b->geom,
                                                      b->keys[0]);
    /* every instance sorted in its own slot range: radix tiles never cross
     * an instance, and their histograms are laid out instance-major, so the
     * scan places every key inside its instance (no instance digit pass) */
    const unsigned nblocks = b->sort_tiles;
    const unsigned total = kRadix * nblocks;
    const unsigned segs = total < kScanBlocks * kBuildBlock ? (total + kBuildBlock - 1) / kBuildBlock : kScanBlocks;
    for (unsigned pass = 0; pass < kRadixPasses; ++pass) {
        const unsigned shift = pass * kRadixBits;
        const unsigned in = pass & 1u, out = in ^ 1u;
        k_radix_hist<<<nblocks, kRadixBlock, 0, s>>>(b->keys[in], n, shift, b->block_hist, nblocks, b->sort_seg,
                                                     b->key_max);
        k_scan_reduce<<<segs, kBuildBlock, 0, s>>>(b->block_hist, total, b->scan_sums, b->key_max, shift);
        k_scan_down<<<segs, kBuildBlock, 0, s>>>(b->block_hist, total, b->scan_sums, b->key_max, shift);
        k_radix_scatter<<<nblocks, kRadixBlock, 0, s>>>(b->keys[in], b->vals[in], b->keys[out], b->vals[out], n,
                                                        shift, b->block_hist, nblocks, pass == 0, b->sort_seg,
                                                        b->key_max);
    }
    /* kRadixPasses is even: the sorted keys and source slots are back in buffer 0 */
    const unsigned f = 0;
    const unsigned long long per_instance = b->max_instance_cells;
    const unsigned clear_x = blocks_for(per_instance, kBuildBlock, 256);
    k_batch_ranges_clear<<<dim3(clear_x, b->instances), kBuildBlock, 0, s>>>(b->ranges, b->inst, b->geom);
    k_batch_ranges<<<agent_blocks, kBuildBlock, 0, s>>>(b->keys[f], n, b->tile_inst, b->inst, b->dims, b->geom,
                                                        b->ranges);
    k_batch_permute<<<agent_blocks, kBuildBlock, 0, s>>>(b->keys[f], b->vals[f], n, b->words, cap, b->tile_inst,
                                                         b->inst, b->dims, b->geom, b->ranges, src, dst);
    const unsigned tile_blocks = blocks_for(b->tiles, kBuildBlock, 0xFFFFFFFFu);
    if (b->order_mode == TAY_GPU_TILE_ORDER_COST) {
        const float *radii = nullptr;
        for (const TayGpuBatch::Pass &p : b->passes)
            if (!radii && p.desc.type == TAY_GPU_PASS_SEE)
                radii = p.radii;
        k_tile_cost<<<tile_blocks, kBuildBlock, 0, s>>>(b->keys[f], b->tiles, b->tile_inst, b->inst, b->dims,
                                                        b->geom, b->ranges, radii, b->tile_bucket, b->bucket_hist);
        k_tile_scan<<<1, 64, 0, s>>>(b->bucket_hist, b->bucket_off, b->queues, (unsigned)b->passes.size());
        k_tile_scatter<<<tile_blocks, kBuildBlock, 0, s>>>(b->tile_bucket, b->tiles, b->bucket_off, b->order);
    } else {
        k_tile_identity<<<tile_blocks, kBuildBlock, 0, s>>>(b->tiles, b->order, b->queues,
                                                            (unsigned)b->passes.size());
    }
    if (!launch_ok("batch sort"))
        return false;
    b->cur ^= 1;

    unsigned *cols = b->cols[b->cur];
    for (unsigned i = 0; i < b->passes.size(); ++i) {
        TayGpuBatch::Pass &p = b->passes[i];
        TayBatchSeeArgs &see = p.see_args[b->cur];
        TayBatchActArgs &act = p.act_args[b->cur];
        void *arg;
        size_t arg_size;
        unsigned grid;
        if (p.desc.type == TAY_GPU_PASS_SEE) {
            std::memset(&see, 0, sizeof(see));
            see.cols = cols;
            see.cap = cap;
            see.tiles = b->tiles;
            see.tile_inst = b->tile_inst;
            see.order = b->order;
            see.queue = b->queues + i;
            see.inst = b->inst;
            see.geom = b->geom;
            see.ranges = b->ranges;
            see.radii = p.radii;
            see.context = p.context;
            see.ctx_words = p.ctx_words;
            see.width = b->width;
            see.lane_walk = b->lane_walk;
            see.pair_count = p.pair_count;
            arg = &see;
            arg_size = sizeof(see);
            grid = p.grid;
        } else {
            act.cols = cols;
            act.cap = cap;
            act.slots = n;
            act.tile_inst = b->tile_inst;
            act.inst = b->inst;
            act.context = p.context;
            act.ctx_words = p.ctx_words;
            arg = &act;
            arg_size = sizeof(act);
            grid = blocks_for(n, kActBlock, 0xFFFFFFFFu);
        }
        void *config[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, arg, HIP_LAUNCH_PARAM_BUFFER_SIZE, &arg_size,
                          HIP_LAUNCH_PARAM_END};
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (record && (!take_event(p, &e0) || !take_event(p, &e1) || !hip_ok(hipEventRecord(e0, s), "hipEventRecord")))
            return false;
        const unsigned block = p.desc.type == TAY_GPU_PASS_SEE ? kSeeBlock : kActBlock;
        if (!hip_ok(hipModuleLaunchKernel(p.kernel, grid, 1, 1, block, 1, 1, 0, s, nullptr, config),
                    "hipModuleLaunchKernel"))
            return false;
        if (record) {
            if (!hip_ok(hipEventRecord(e1, s), "hipEventRecord"))
                return false;
            p.ev_pending.emplace_back(e0, e1);
            ++p.launches;
        }
    }
    return true;
}

}  // namespace

/* -------------------------------------------------------------- C-ABI */

TAY_API void tay_gpu_batch_destroy(TayGpuBatch *b) {
    if (!b)
        return;
    (void)tay_gpu_synchronize();
    drop_graph(b);
    for (auto &p : b->passes) {
        device_free(p.context);
        device_free(p.radii);
        device_free(p.pair_count);
        (void)resolve_timing(p);
        for (hipEvent_t e : p.ev_free)
            (void)hipEventDestroy(e);
    }
    void *bufs[] = {b->cols[0], b->cols[1], b->inst, b->tile_inst, b->parts, b->boxes, b->geom, b->keys[0],
                    b->keys[1], b->vals[0], b->vals[1], b->block_hist, b->scan_sums, b->ranges, b->tile_bucket,
                    b->bucket_hist, b->bucket_off, b->order, b->queues, b->staging, b->backup, b->sort_seg};
    for (void *p : bufs)
        device_free(p);
    if (b->module)
        (void)hipModuleUnload(b->module);
    delete b;
}

TAY_API TayGpuBatch *tay_gpu_batch_create(const char *const *sources, unsigned sources_count,
                                          const TayGpuBatchPass *passes, unsigned passes_count, unsigned agent_size,
                                          int dims, unsigned instances, const unsigned *counts,
                                          const unsigned long long *cell_table_bytes, int count_pairs, char *log,
                                          unsigned log_size) {
    if (instances == 0 || instances > 65535) {
        set_error("a batch holds 1 to 65535 instances (%u given)", instances);
        return nullptr;
    }
    if (dims < 1 || dims > 3) {
        set_error("%d dimensions not supported by the batched Z-grid (1 to 3)", dims);
        return nullptr;
    }
    if (agent_size % 4 != 0 || agent_size < 8u + 4u * (unsigned)dims) {
        set_error("agent size %u must be a multiple of 4 of at least %u bytes (tag and position)", agent_size,
                  8u + 4u * (unsigned)dims);
        return nullptr;
    }
    if (passes_count == 0 || passes_count > 64) {
        set_error("a batch runs 1 to 64 passes (%u given)", passes_count);
        return nullptr;
    }
    for (unsigned i = 0; i < passes_count; ++i)
        if (!valid_func_name(passes[i].func_name) ||
            (passes[i].type != TAY_GPU_PASS_SEE && passes[i].type != TAY_GPU_PASS_ACT)) {
            set_error("pass %u: function name is not a C identifier or the type is unknown", i);
            return nullptr;
        }
    unsigned long long slots = 0, records = 0, cells = 0, max_cells = 0;
    std::vector<TayBatchInstance> inst(instances);
    for (unsigned b = 0; b < instances; ++b) {
        const unsigned long long c = cell_table_bytes[b] / sizeof(TayCellRange);
        if (c == 0) {
            set_error("instance %u: cell table needs at least %zu bytes", b, sizeof(TayCellRange));
            return nullptr;
        }
        inst[b].off = (unsigned)slots;
        inst[b].count = counts[b];
        inst[b].rec = (unsigned)records;
        inst[b].pad = 0;
        inst[b].ranges_off = cells;
        inst[b].max_cells = c;
        slots += ((unsigned long long)counts[b] + 63ull) & ~63ull;
        records += counts[b];
        cells += c;
        max_cells = c > max_cells ? c : max_cells;
    }
    if (slots >= 0xFFFFFFFFull - 64u || slots == 0) {
        set_error("batch of %llu slots: needs between 1 and 2^32 - 65", slots);
        return nullptr;
    }
    /* kernels index a column word as word * slots + slot, and a slot's byte
     * offset in a column, in 32 bits */
    if (slots * (agent_size / 4u) > 0xFFFFFFFFull || slots > (1ull << 30)) {
        set_error("batch of %llu slots of %u-byte agents exceeds 2^32 column words or 2^30 slots", slots,
                  agent_size);
        return nullptr;
    }
    if (!stream())
        return nullptr;

    std::vector<char> code;
    if (!compile_program(compose_batch(sources, sources_count, passes, passes_count, agent_size, dims, count_pairs),
                         code, log, log_size))
        return nullptr;

    TayGpuBatch *b = new TayGpuBatch();
    b->agent_size = agent_size;
    b->words = agent_size / 4;
    b->instances = instances;
    b->records = (unsigned)records;
    b->slots = (unsigned)slots;
    b->tiles = (unsigned)(slots / 64);
    b->dims = dims;
    b->count_pairs = count_pairs;
    b->timing = 0;
    b->width = 64;
    b->order_mode = TAY_GPU_TILE_ORDER_COST;
    b->order_bits = TAY_GPU_DEFAULT_ORDER_BITS;
    b->lane_walk = TAY_GPU_DEFAULT_LANE_WALK;
    b->inst_host = inst;
    b->radius_host.assign(4ull * instances, 0.0f);
    b->total_cells = cells;
    b->max_instance_cells = max_cells;
    b->cols[0] = b->cols[1] = nullptr;
    b->cur = 0;
    b->inst = nullptr;
    b->tile_inst = nullptr;
    b->parts = b->boxes = nullptr;
    b->geom = nullptr;
    b->keys[0] = b->keys[1] = b->vals[0] = b->vals[1] = nullptr;
    b->block_hist = b->scan_sums = b->key_max = nullptr;
    b->sort_seg = nullptr;
    b->ranges = nullptr;
    b->tile_bucket = b->bucket_hist = b->bucket_off = b->order = b->queues = nullptr;
    b->staging = b->backup = nullptr;
    b->module = nullptr;
    b->graph[0] = b->graph[1] = nullptr;
    b->exec[0] = b->exec[1] = nullptr;

    hipStream_t s = stream();
    const size_t col_bytes = (size_t)b->words * slots * 4;
    std::vector<TaySortTile> seg;
    unsigned long long hist_words = 0;
    for (unsigned i = 0; i < instances; ++i) {
        const unsigned padded = (inst[i].count + 63u) & ~63u;
        const unsigned nb = (padded + kRadixTile - 1) / kRadixTile;
        for (unsigned k = 0; k < nb; ++k) {
            TaySortTile t;
            t.start = inst[i].off + k * kRadixTile;
            t.count = padded - k * kRadixTile < kRadixTile ? padded - k * kRadixTile : kRadixTile;
            t.hist_base = (unsigned)hist_words + k;
            t.hist_stride = nb;
            seg.push_back(t);
        }
        hist_words += (unsigned long long)kRadix * nb;
    }
    b->sort_tiles = (unsigned)seg.size();
    const unsigned nblocks = b->sort_tiles ? b->sort_tiles : 1u;
    std::vector<unsigned> tile_inst(b->tiles);
    for (unsigned i = 0; i < instances; ++i)
        for (unsigned t = inst[i].off / 64u; t < (inst[i].off + inst[i].count + 63u) / 64u; ++t)
            tile_inst[t] = i;
    std::vector<float> parts(4ull * instances, 1.0f);
    b->parts_host = parts;
    bool ok = device_alloc((void **)&b->cols[0], col_bytes, "batch columns") &&
              device_alloc((void **)&b->cols[1], col_bytes, "batch columns") &&
              hip_ok(hipMemsetAsync(b->cols[0], 0, col_bytes, s), "hipMemsetAsync") &&
              hip_ok(hipMemsetAsync(b->cols[1], 0, col_bytes, s), "hipMemsetAsync") &&
              device_alloc((void **)&b->inst, sizeof(TayBatchInstance) * instances, "batch instances") &&
              device_alloc((void **)&b->tile_inst, 4ull * b->tiles, "tile instances") &&
              device_alloc((void **)&b->parts, 16ull * instances, "part sizes") &&
              device_alloc((void **)&b->boxes, 24ull * b->tiles, "tile boxes") &&
              device_alloc((void **)&b->geom, sizeof(TayGridGeom) * instances, "geometry") &&
              device_alloc((void **)&b->keys[0], 4ull * slots, "sort keys") &&
              device_alloc((void **)&b->keys[1], 4ull * slots, "sort keys") &&
              device_alloc((void **)&b->vals[0], 4ull * slots, "sort values") &&
              device_alloc((void **)&b->vals[1], 4ull * slots, "sort values") &&
              device_alloc((void **)&b->block_hist, 4ull * kRadix * nblocks, "radix histogram") &&
              device_alloc((void **)&b->sort_seg, sizeof(TaySortTile) * nblocks, "sort tiles") &&
              device_alloc((void **)&b->scan_sums, 4ull * (kScanBlocks + 1), "scan sums") &&
              device_alloc((void **)&b->ranges, sizeof(TayCellRange) * cells, "cell tables") &&
              device_alloc((void **)&b->tile_bucket, 4ull * b->tiles, "tile buckets") &&
              device_alloc((void **)&b->bucket_hist, 4ull * kCostBuckets, "bucket histogram") &&
              device_alloc((void **)&b->bucket_off, 4ull * kCostBuckets, "bucket offsets") &&
              device_alloc((void **)&b->order, 4ull * b->tiles, "tile order") &&
              device_alloc((void **)&b->queues, 4ull * passes_count, "work queues") &&
              hip_ok(hipMemsetAsync(b->bucket_hist, 0, 4ull * kCostBuckets, s), "hipMemsetAsync") &&
              hip_ok(hipMemsetAsync(b->geom, 0, sizeof(TayGridGeom) * instances, s), "hipMemsetAsync") &&
              hip_ok(hipMemcpyAsync(b->inst, inst.data(), sizeof(TayBatchInstance) * instances,
                                    hipMemcpyHostToDevice, s), "upload instances") &&
              hip_ok(hipMemcpyAsync(b->tile_inst, tile_inst.data(), 4ull * b->tiles, hipMemcpyHostToDevice, s),
                     "upload tile instances") &&
              (seg.empty() || hip_ok(hipMemcpyAsync(b->sort_seg, seg.data(), sizeof(TaySortTile) * seg.size(),
                                                    hipMemcpyHostToDevice, s), "upload sort tiles")) &&
              hip_ok(hipMemcpyAsync(b->parts, parts.data(), 16ull * instances, hipMemcpyHostToDevice, s),
                     "upload part sizes") &&
              hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize") &&
              hip_ok(hipModuleLoadData(&b->module, code.data()), "hipModuleLoadData");
    if (ok)
        b->key_max = b->scan_sums + kScanBlocks;
    int cus = 0;
    ok = ok && hip_ok(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0), "hipDeviceGetAttribute");
    b->passes.resize(passes_count);
    for (unsigned i = 0; ok && i < passes_count; ++i) {
        TayGpuBatch::Pass &p = b->passes[i];
        p.desc = passes[i];
        p.func_name = passes[i].func_name;
        p.desc.func_name = p.func_name.c_str();
        p.ctx_words = (passes[i].context_size + 3u) / 4u;
        p.kernel = nullptr;
        p.context = nullptr;
        p.radii = nullptr;
        p.pair_count = nullptr;
        p.grid = 0;
        p.ms = 0.0;
        p.launches = 0;
        char name[64];
        std::snprintf(name, sizeof(name), "tay_batch_pass_%u", i);
        const size_t ctx_bytes = 4ull * (p.ctx_words ? p.ctx_words : 1u) * instances;
        ok = hip_ok(hipModuleGetFunction(&p.kernel, b->module, name), "hipModuleGetFunction") &&
             device_alloc((void **)&p.context, ctx_bytes, "batch contexts") &&
             hip_ok(hipMemsetAsync(p.context, 0, ctx_bytes, s), "hipMemsetAsync");
        if (ok && p.desc.type == TAY_GPU_PASS_SEE) {
            int per_cu = 0;
            ok = device_alloc((void **)&p.radii, 16ull * instances, "batch radii") &&
                 hip_ok(hipMemsetAsync(p.radii, 0, 16ull * instances, s), "hipMemsetAsync") &&
                 device_alloc((void **)&p.pair_count, 8ull * TAY_PAIR_SLOTS, "pair counters") &&
                 hip_ok(hipMemsetAsync(p.pair_count, 0, 8ull * TAY_PAIR_SLOTS, s), "hipMemsetAsync") &&
                 hip_ok(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, p.kernel, kSeeBlock, 0),
                        "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor");
            /* tuning: TAY_BATCH_SEE_BLOCKS_PER_CU=<n> replaces the occupancy
             * query's resident blocks per compute unit */
            if (const char *e = std::getenv("TAY_BATCH_SEE_BLOCKS_PER_CU")) {
                const int n = std::atoi(e);
                if (n >= 1 && n <= 32)
                    per_cu = n;
            }
            if (std::getenv("TAY_BATCH_VERBOSE"))
                std::fprintf(stderr, "batch see pass %u: %d blocks per CU, %d CUs\n", i, per_cu, cus);
            const unsigned want = (unsigned)(per_cu > 0 ? per_cu : 1) * (unsigned)(cus > 0 ? cus : 1);
            const unsigned need = (b->tiles + kSeeBlock / 64u - 1u) / (kSeeBlock / 64u);
            p.grid = want < need ? want : need;
        }
    }
    if (!ok) {
        tay_gpu_batch_destroy(b);
        return nullptr;
    }
    return b;
}

TAY_API int tay_gpu_batch_compile_check(const char *const *sources, unsigned sources_count,
                                        const TayGpuBatchPass *passes, unsigned passes_count, unsigned agent_size,
                                        int dims, char *log, unsigned log_size, unsigned long long *code_size) {
    for (unsigned i = 0; i < passes_count; ++i)
        if (!valid_func_name(passes[i].func_name)) {
            set_error("pass %u: function name is not a C identifier", i);
            return 0;
        }
    if (agent_size % 4 != 0 || dims < 1 || dims > 3) {
        set_error("agent size must be a multiple of 4 and dims 1 to 3");
        return 0;
    }
    std::vector<char> code;
    if (!compile_program(compose_batch(sources, sources_count, passes, passes_count, agent_size, dims, 1), code, log,
                         log_size))
        return 0;
    if (code_size)
        *code_size = code.size();
    return 1;
}

TAY_API int tay_gpu_batch_set_part_sizes(TayGpuBatch *b, const float *part_sizes) {
    for (unsigned i = 0; i < b->instances; ++i)
        for (int k = 0; k < b->dims; ++k)
            if (!(part_sizes[4ull * i + k] > 0.0f)) {
                set_error("instance %u: partition size of axis %d must be positive", i, k);
                return 0;
            }
    hipStream_t s = stream();
    if (!(s && hip_ok(hipMemcpyAsync(b->parts, part_sizes, 16ull * b->instances, hipMemcpyHostToDevice, s),
                      "upload part sizes") &&
          hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize")))
        return 0;
    b->parts_host.assign(part_sizes, part_sizes + 4ull * b->instances);
    return 1;
}

TAY_API int tay_gpu_batch_get_part_sizes(const TayGpuBatch *b, float *part_sizes) {
    std::memcpy(part_sizes, b->parts_host.data(), 16ull * b->instances);
    return 1;
}

TAY_API int tay_gpu_batch_set_radii(TayGpuBatch *b, unsigned pass_index, const float *radii) {
    if (pass_index >= b->passes.size() || b->passes[pass_index].desc.type != TAY_GPU_PASS_SEE) {
        set_error("pass %u is not a see pass", pass_index);
        return 0;
    }
    bool first_see = true;
    for (unsigned i = 0; i < pass_index; ++i)
        first_see = first_see && b->passes[i].desc.type != TAY_GPU_PASS_SEE;
    if (first_see)
        b->radius_host.assign(radii, radii + 4ull * b->instances);
    hipStream_t s = stream();
    return s && hip_ok(hipMemcpyAsync(b->passes[pass_index].radii, radii, 16ull * b->instances,
                                      hipMemcpyHostToDevice, s), "upload radii") &&
           hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");
}

TAY_API int tay_gpu_batch_set_contexts(TayGpuBatch *b, unsigned pass_index, const void *contexts) {
    if (pass_index >= b->passes.size()) {
        set_error("pass index %u out of range", pass_index);
        return 0;
    }
    TayGpuBatch::Pass &p = b->passes[pass_index];
    if (!p.desc.context_size)
        return 1;
    /* one instance per ctx_words words, so every context starts word-aligned */
    std::vector<unsigned> words((size_t)p.ctx_words * b->instances, 0u);
    for (unsigned i = 0; i < b->instances; ++i)
        std::memcpy(&words[(size_t)i * p.ctx_words], (const char *)contexts + (size_t)i * p.desc.context_size,
                    p.desc.context_size);
    hipStream_t s = stream();
    return s && hip_ok(hipMemcpyAsync(p.context, words.data(), 4ull * words.size(), hipMemcpyHostToDevice, s),
                       "upload contexts") &&
           hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");
}

TAY_API int tay_gpu_batch_set_tile_width(TayGpuBatch *b, unsigned width) {
    if (width != 64 && width != 32 && width != 16 && width != 8) {
        set_error("tile width %u: must be 64, 32, 16 or 8", width);
        return 0;
    }
    if (width != b->width)
        drop_graph(b);
    b->width = width;
    return 1;
}

TAY_API int tay_gpu_batch_set_tile_order(TayGpuBatch *b, int order) {
    if (order != TAY_GPU_TILE_ORDER_COST && order != TAY_GPU_TILE_ORDER_STORAGE) {
        set_error("unknown tile order %d", order);
        return 0;
    }
    if (order != b->order_mode)
        drop_graph(b);
    b->order_mode = order;
    return 1;
}

TAY_API int tay_gpu_batch_set_lane_walk(TayGpuBatch *b, unsigned width) {
    if (width > 64) {
        set_error("lane walk width %u: must be at most 64", width);
        return 0;
    }
    if (width != b->lane_walk)
        drop_graph(b);
    b->lane_walk = width;
    return 1;
}

TAY_API int tay_gpu_batch_set_order_bits(TayGpuBatch *b, int bits) {
    if (bits < 0 || bits > 3) {
        set_error("order bits %d: must be 0 to 3", bits);
        return 0;
    }
    if (bits != b->order_bits)
        drop_graph(b);
    b->order_bits = bits;
    return 1;
}

static bool batch_staging(TayGpuBatch *b) {
    return b->staging || device_alloc((void **)&b->staging, (size_t)b->records * b->agent_size + 4, "batch staging");
}

TAY_API int tay_gpu_batch_push(TayGpuBatch *b, const void *agents) {
    hipStream_t s = stream();
    if (!s || !batch_staging(b))
        return 0;
    if (b->records == 0)
        return 1;
    const size_t bytes = (size_t)b->records * b->agent_size;
    if (!hip_ok(hipMemcpyAsync(b->staging, agents, bytes, hipMemcpyHostToDevice, s), "push agents"))
        return 0;
    k_batch_transpose<<<blocks_for((unsigned long long)b->records * b->words, kBuildBlock, 8192), kBuildBlock, 0,
                        s>>>(b->staging, b->cols[b->cur], b->records, b->words, b->slots, b->inst, b->instances, 1);
    return launch_ok("batch push") && hip_ok(hipStreamSynchronize(s), "push agents");
}

TAY_API int tay_gpu_batch_fetch(TayGpuBatch *b, void *agents) {
    hipStream_t s = stream();
    if (!s || !batch_staging(b))
        return 0;
    if (b->records == 0)
        return 1;
    k_batch_transpose<<<blocks_for((unsigned long long)b->records * b->words, kBuildBlock, 8192), kBuildBlock, 0,
                        s>>>(b->staging, b->cols[b->cur], b->records, b->words, b->slots, b->inst, b->instances, 0);
    return launch_ok("batch fetch") &&
           hip_ok(hipMemcpyAsync(agents, b->staging, (size_t)b->records * b->agent_size, hipMemcpyDeviceToHost, s),
                  "fetch agents") &&
           hip_ok(hipStreamSynchronize(s), "fetch agents");
}

TAY_API int tay_gpu_batch_run(TayGpuBatch *b, unsigned steps, int use_graph, double *device_ms) {
    hipStream_t s = stream();
    if (!s)
        return 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (device_ms && (!hip_ok(hipEventCreate(&e0), "hipEventCreate") || !hip_ok(hipEventCreate(&e1), "hipEventCreate") ||
                      !hip_ok(hipEventRecord(e0, s), "hipEventRecord")))
        return 0;
    bool ok = true;
    if (use_graph && steps >= 2) {
        /* Two steps return the columns to the buffer they started from, so a
         * captured pair of steps replays for as long as needed. */
        const int p = b->cur;
        if (!b->exec[p]) {
            ok = hip_ok(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            if (ok) {
                const bool first = batch_step(b, s, false);
                const bool second = first && batch_step(b, s, false);
                hipGraph_t g = nullptr;
                ok = hip_ok(hipStreamEndCapture(s, &g), "hipStreamEndCapture") && first && second;
                b->graph[p] = g;
                ok = ok && hip_ok(hipGraphInstantiate(&b->exec[p], g, nullptr, nullptr, 0), "hipGraphInstantiate");
                b->cur = p;
                if (!ok)
                    drop_graph(b);
            }
        }
        for (unsigned k = 0; ok && k + 1 < steps; k += 2)
            ok = hip_ok(hipGraphLaunch(b->exec[p], s), "hipGraphLaunch");
        if (ok && (steps & 1u))
            ok = batch_step(b, s, false);
    } else {
        for (unsigned k = 0; ok && k < steps; ++k)
            ok = batch_step(b, s, b->timing && !use_graph);
    }
    if (device_ms) {
        float ms = 0.0f;
        ok = ok && hip_ok(hipEventRecord(e1, s), "hipEventRecord") &&
             hip_ok(hipEventSynchronize(e1), "hipEventSynchronize") &&
             hip_ok(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        *device_ms = ms;
    }
    if (e0)
        (void)hipEventDestroy(e0);
    if (e1)
        (void)hipEventDestroy(e1);
    return ok;
}

TAY_API int tay_gpu_batch_enable_timing(TayGpuBatch *b, int enable) {
    bool ok = true;
    for (auto &p : b->passes) {
        ok = resolve_timing(p) && ok;
        p.ms = 0.0;
        p.launches = 0;
    }
    b->timing = enable;
    return ok;
}

TAY_API int tay_gpu_batch_pass_time(TayGpuBatch *b, unsigned pass_index, double *ms, unsigned *launches) {
    if (pass_index >= b->passes.size()) {
        set_error("pass index %u out of range", pass_index);
        return 0;
    }
    TayGpuBatch::Pass &p = b->passes[pass_index];
    if (!resolve_timing(p))
        return 0;
    *ms = p.ms;
    *launches = p.launches;
    return 1;
}

TAY_API int tay_gpu_batch_pair_count(TayGpuBatch *b, unsigned pass_index, unsigned long long *pairs) {
    if (pass_index >= b->passes.size() || !b->passes[pass_index].pair_count) {
        set_error("pass %u is not a see pass", pass_index);
        return 0;
    }
    if (!b->count_pairs) {
        set_error("batch was created without pair counting");
        return 0;
    }
    return read_pair_counters(b->passes[pass_index].pair_count, pairs);
}

TAY_API int tay_gpu_batch_reset_pair_counts(TayGpuBatch *b) {
    hipStream_t s = stream();
    if (!s)
        return 0;
    for (auto &p : b->passes)
        if (p.pair_count && !hip_ok(hipMemsetAsync(p.pair_count, 0, 8ull * TAY_PAIR_SLOTS, s), "reset pair count"))
            return 0;
    return 1;
}

namespace {

bool radii_set(const TayGpuBatch *b, const char *what) {
    for (unsigned i = 0; i < b->instances; ++i)
        for (int k = 0; k < b->dims; ++k)
            if (!(b->radius_host[4ull * i + k] > 0.0f)) {
                set_error("%s: instance %u has no positive radius on axis %d (set the first see pass's radii)", what,
                          i, k);
                return false;
            }
    return true;
}

bool tune_widths_ok(const unsigned *widths, unsigned widths_count, const unsigned *lane_walks,
                    unsigned lane_walks_count, unsigned steps) {
    if (!widths_count || !lane_walks_count || !steps) {
        set_error("autotune needs at least one tile width, lane-walk width and step");
        return false;
    }
    for (unsigned j = 0; j < lane_walks_count; ++j)
        if (lane_walks[j] > 64) {
            set_error("lane walk width %u: must be at most 64", lane_walks[j]);
            return false;
        }
    for (unsigned j = 0; j < widths_count; ++j)
        if (widths[j] != 64 && widths[j] != 32 && widths[j] != 16 && widths[j] != 8) {
            set_error("tile width %u: must be 64, 32, 16 or 8", widths[j]);
            return false;
        }
    return true;
}

/* Every instance's geometry, and so the bounding box of its agents, from
 * the current columns and part sizes. */
bool instance_geometry(TayGpuBatch *b, std::vector<TayGridGeom> &geom) {
    hipStream_t s = stream();
    if (!s)
        return false;
    if (b->tiles)
        k_tile_box<<<(b->tiles + kBuildBlock / 64u - 1u) / (kBuildBlock / 64u), kBuildBlock, 0, s>>>(
            b->cols[b->cur], b->slots, b->tiles, b->tile_inst, b->inst, b->dims, b->boxes, b->key_max);
    k_batch_geometry<<<b->instances, 64, 0, s>>>(b->boxes, b->parts, b->inst, b->dims, b->order_bits, b->geom,
                                                 b->key_max);
    geom.resize(b->instances);
    return launch_ok("instance geometry") &&
           hip_ok(hipMemcpyAsync(geom.data(), b->geom, sizeof(TayGridGeom) * b->instances, hipMemcpyDeviceToHost,
                                 s),
                  "download geometry") &&
           hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");
}

/* Cell sizes for `target` agents per cell: an instance of n agents whose
 * bounding box has volume V (area, length in fewer dimensions) gets cells of
 * side (target V / n)^(1/dims), clamped per axis to [min_factor, max_factor]
 * times its radius. */
void density_parts(const TayGpuBatch *b, const std::vector<TayGridGeom> &geom, float target, float min_factor,
                   float max_factor, float *parts) {
    for (unsigned i = 0; i < b->instances; ++i) {
        double vol = 1.0;
        for (int k = 0; k < b->dims; ++k)
            vol *= std::max(0.0, (double)geom[i].box_max[k] - (double)geom[i].box_min[k]);
        const unsigned n = b->inst_host[i].count;
        const double side = n && vol > 0.0 ? std::pow((double)target * vol / n, 1.0 / b->dims) : 0.0;
        for (int k = 0; k < 4; ++k) {
            float part = 1.0f;
            if (k < b->dims) {
                const double r = b->radius_host[4ull * i + k];
                part = (float)std::min(std::max(side, (double)min_factor * r), (double)max_factor * r);
            }
            parts[4ull * i + k] = part;
        }
    }
}

bool density_args_ok(float min_factor, float max_factor) {
    if (!(min_factor > 0.0f && min_factor <= max_factor && std::isfinite(max_factor))) {
        set_error("cell factor bounds [%g, %g]: need 0 < min <= max", min_factor, max_factor);
        return false;
    }
    return true;
}

/* Times every candidate set of part sizes (4 per instance each) with every
 * tile width and lane-walk width, restoring the starting agents before each
 * run of `steps` un-captured steps; table_ms (optional, candidates x widths x
 * lane walks, lane walk fastest) receives ms per step.  Applies the fastest
 * combination and leaves the agents as they were. */
bool autotune_parts(TayGpuBatch *b, const std::vector<std::vector<float>> &cands, const unsigned *widths,
                    unsigned widths_count, const unsigned *lane_walks, unsigned lane_walks_count, unsigned steps,
                    double *table_ms, unsigned *best_cand, unsigned *best_width, unsigned *best_lane_walk) {
    hipStream_t s = stream();
    const size_t col_bytes = (size_t)b->words * b->slots * 4;
    if (!s || (!b->backup && !device_alloc((void **)&b->backup, col_bytes, "autotune backup")))
        return false;
    const int start_cur = b->cur;
    const int timing = b->timing;
    b->timing = 0;
    bool ok = hip_ok(hipMemcpyAsync(b->backup, b->cols[start_cur], col_bytes, hipMemcpyDeviceToDevice, s), "backup");
    double best = 1e300;
    unsigned bc = 0, bw = widths[0], bl = lane_walks[0];
    for (unsigned i = 0; ok && i < cands.size(); ++i) {
        ok = tay_gpu_batch_set_part_sizes(b, cands[i].data());
        for (unsigned j = 0; ok && j < widths_count; ++j)
            for (unsigned l = 0; ok && l < lane_walks_count; ++l) {
                b->width = widths[j];
                b->lane_walk = lane_walks[l];
                b->cur = start_cur;
                double ms = 0.0;
                ok = hip_ok(hipMemcpyAsync(b->cols[start_cur], b->backup, col_bytes, hipMemcpyDeviceToDevice, s),
                            "restore") &&
                     tay_gpu_batch_run(b, steps, 0, &ms);
                ms /= steps;
                if (table_ms)
                    table_ms[((size_t)i * widths_count + j) * lane_walks_count + l] = ms;
                if (ok && ms < best) {
                    best = ms;
                    bc = i;
                    bw = widths[j];
                    bl = lane_walks[l];
                }
            }
    }
    b->cur = start_cur;
    ok = ok && hip_ok(hipMemcpyAsync(b->cols[start_cur], b->backup, col_bytes, hipMemcpyDeviceToDevice, s), "restore");
    ok = ok && tay_gpu_batch_set_part_sizes(b, cands[bc].data()) && tay_gpu_batch_reset_pair_counts(b);
    drop_graph(b);
    b->width = bw;
    b->lane_walk = bl;
    b->timing = timing;
    *best_cand = bc;
    if (best_width)
        *best_width = bw;
    if (best_lane_walk)
        *best_lane_walk = bl;
    return ok;
}

}  // namespace

TAY_API int tay_gpu_batch_autotune(TayGpuBatch *b, const float *cell_factors, unsigned factors_count,
                                   const unsigned *widths, unsigned widths_count, const unsigned *lane_walks,
                                   unsigned lane_walks_count, unsigned steps, double *table_ms, float *best_factor,
                                   unsigned *best_width, unsigned *best_lane_walk) {
    if (!factors_count) {
        set_error("autotune needs at least one cell factor, tile width, lane-walk width and step");
        return 0;
    }
    if (!tune_widths_ok(widths, widths_count, lane_walks, lane_walks_count, steps))
        return 0;
    for (unsigned i = 0; i < factors_count; ++i)
        if (!(cell_factors[i] > 0.0f)) {
            set_error("cell factor %u must be positive", i);
            return 0;
        }
    if (!radii_set(b, "autotune"))
        return 0;
    std::vector<std::vector<float>> cands(factors_count, std::vector<float>(4ull * b->instances, 1.0f));
    for (unsigned i = 0; i < factors_count; ++i)
        for (unsigned k = 0; k < b->instances; ++k)
            for (int d = 0; d < b->dims; ++d)
                cands[i][4ull * k + d] = cell_factors[i] * b->radius_host[4ull * k + d];
    unsigned bc = 0;
    const bool ok = autotune_parts(b, cands, widths, widths_count, lane_walks, lane_walks_count, steps, table_ms, &bc,
                                   best_width, best_lane_walk);
    if (best_factor)
        *best_factor = cell_factors[bc];
    return ok;
}

TAY_API int tay_gpu_batch_density_part_sizes(TayGpuBatch *b, float target, float min_factor, float max_factor,
                                             float *part_sizes) {
    if (!(target > 0.0f && std::isfinite(target))) {
        set_error("agents per cell %g: must be positive", target);
        return 0;
    }
    std::vector<TayGridGeom> geom;
    if (!density_args_ok(min_factor, max_factor) || !radii_set(b, "density part sizes") ||
        !instance_geometry(b, geom))
        return 0;
    density_parts(b, geom, target, min_factor, max_factor, part_sizes);
    return 1;
}

TAY_API int tay_gpu_batch_autotune_density(TayGpuBatch *b, const float *targets, unsigned targets_count,
                                           float min_factor, float max_factor, const unsigned *widths,
                                           unsigned widths_count, const unsigned *lane_walks,
                                           unsigned lane_walks_count, unsigned steps, double *table_ms,
                                           float *best_target, float *part_sizes, unsigned *best_width,
                                           unsigned *best_lane_walk) {
    if (!targets_count) {
        set_error("autotune needs at least one agents-per-cell target");
        return 0;
    }
    for (unsigned i = 0; i < targets_count; ++i)
        if (!(targets[i] > 0.0f && std::isfinite(targets[i]))) {
            set_error("agents per cell %u must be positive", i);
            return 0;
        }
    std::vector<TayGridGeom> geom;
    if (!tune_widths_ok(widths, widths_count, lane_walks, lane_walks_count, steps) ||
        !density_args_ok(min_factor, max_factor) || !radii_set(b, "autotune") || !instance_geometry(b, geom))
        return 0;
    std::vector<std::vector<float>> cands(targets_count, std::vector<float>(4ull * b->instances, 1.0f));
    for (unsigned i = 0; i < targets_count; ++i)
        density_parts(b, geom, targets[i], min_factor, max_factor, cands[i].data());
    unsigned bc = 0;
    const bool ok = autotune_parts(b, cands, widths, widths_count, lane_walks, lane_walks_count, steps, table_ms, &bc,
                                   best_width, best_lane_walk);
    if (best_target)
        *best_target = targets[bc];
    if (part_sizes)
        std::memcpy(part_sizes, cands[bc].data(), 16ull * b->instances);
    return ok;
}
