/*
 * grid_build.h - device code of the Z-grid build shared by the single-space
 * build (gpu_space.hip) and the batched ensembles (gpu_batch.hip): Morton
 * keys, grid geometry, the stable LSD radix sort, cell ranges, the column
 * permutation and the AoS <-> SoA transposes.  Included once per translation
 * unit; everything is internal to it.
 */
#ifndef TAY_GRID_BUILD_H
#define TAY_GRID_BUILD_H

#include "internal.h"

#include <cfloat>
#include <cmath>

namespace {

constexpr unsigned kBuildBlock = 256;
constexpr unsigned kRadixBits = 8;
constexpr unsigned kRadix = 1u << kRadixBits;
constexpr unsigned kRadixPasses = 4;            /* 32-bit keys */
constexpr unsigned kRadixTile = 4096;              /* keys per tile (one block) */
constexpr unsigned kRadixBlock = 1024;             /* threads of the histogram and scatter blocks */
constexpr unsigned kRadixWaves = kRadixBlock / 64;
constexpr unsigned kRadixRounds = kRadixTile / kRadixBlock;   /* 64-key rounds per wavefront */
constexpr unsigned kScanBlocks = 64;             /* segments of the histogram scan, at most one wavefront's lanes */
/* A space of at most kSmallSort agents is sorted by one block (k_small_sort),
 * kSmallBits per pass: keys and source indices (8 bytes per key) and the
 * per-wave digit counts (16 x 512 words) fill the 160 KiB of LDS. */
constexpr unsigned kSmallPerLane = 15;                          /* keys held by one lane */
constexpr unsigned kSmallSort = kSmallPerLane * kRadixBlock;    /* 15,360 */
constexpr unsigned kSmallBits = 9;
constexpr unsigned kSmallRadix = 1u << kSmallBits;

__device__ __forceinline__ unsigned compact2(unsigned v) {
    v &= 0x55555555u;
    v = (v | (v >> 1)) & 0x33333333u;
    v = (v | (v >> 2)) & 0x0F0F0F0Fu;
    v = (v | (v >> 4)) & 0x00FF00FFu;
    v = (v | (v >> 8)) & 0x0000FFFFu;
    return v;
}

__device__ __forceinline__ unsigned compact3(unsigned v) {
    v &= 0x09249249u;
    v = (v | (v >> 2)) & 0x030C30C3u;
    v = (v | (v >> 4)) & 0x0300F00Fu;
    v = (v | (v >> 8)) & 0xFF0000FFu;
    v = (v | (v >> 16)) & 0x000003FFu;
    return v;
}

__device__ __forceinline__ unsigned spread2(unsigned v) {
    v &= 0x0000FFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

__device__ __forceinline__ unsigned spread3(unsigned v) {
    v &= 0x000003FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

/* Morton key of cell indices: 3D 10 bits per axis, 2D 16, 1D identity
 * (tay/cpu_z_grid.c:37-83). */
__device__ __forceinline__ unsigned morton(int x, int y, int z, int dims) {
    if (dims == 3)
        return spread3((unsigned)x) | (spread3((unsigned)y) << 1) | (spread3((unsigned)z) << 2);
    if (dims == 2)
        return spread2((unsigned)x) | (spread2((unsigned)y) << 1);
    return (unsigned)x;
}

/* Cell part of a storage-order key: the key interleaves (cell << s) | sub on
 * every axis, so the cell's Morton key is the key without its low dims x s
 * bits.  Every key is shifted, 0xFFFFFFFF included: a 2D grid of 8192 cells
 * per axis with 3 sub-cell bits gives that key to the agent in the top
 * corner sub-cell, and a batch's padding key (also 0xFFFFFFFF) never reaches
 * this function (its callers skip padding slots). */
__device__ __forceinline__ unsigned cell_of_key(unsigned key, int dims, const TayGridGeom *g) {
    return key >> ((unsigned)dims * (unsigned)g->order_bits);
}

/* Row-major cell id (x fastest) of a cell's Morton key. */
__device__ __forceinline__ unsigned long long morton_to_cell(unsigned key, int dims, const TayGridGeom *g) {
    unsigned x, y = 0, z = 0;
    if (dims == 3) {
        x = compact3(key);
        y = compact3(key >> 1);
        z = compact3(key >> 2);
    } else if (dims == 2) {
        x = compact2(key);
        y = compact2(key >> 1);
    } else {
        x = key;
    }
    return (unsigned long long)x + (unsigned long long)y * (unsigned)g->count[0] +
           (unsigned long long)z * ((unsigned long long)(unsigned)g->count[0] * (unsigned)g->count[1]);
}

/* Row-major cell id of a storage-order key. */
__device__ __forceinline__ unsigned long long key_to_cell(unsigned key, int dims, const TayGridGeom *g) {
    return morton_to_cell(cell_of_key(key, dims, g), dims, g);
}

__device__ __forceinline__ float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1)
        v = fminf(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1)
        v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

/* Per-block bounding box of the positions (min corner for every agent; for
 * non-point agents the max corner at words 6..8) and, for non-point agents,
 * the largest extent max - min per axis: kBoxPartial floats per block. */
constexpr unsigned kBoxPartial = 9;

__global__ __launch_bounds__(kBuildBlock) void k_box_partial(const unsigned *__restrict__ cols, unsigned cap,
                                                             unsigned n, int dims, int is_point,
                                                             float *__restrict__ partials) {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    float ext[3] = {0.0f, 0.0f, 0.0f};
    for (unsigned i = blockIdx.x * kBuildBlock + threadIdx.x; i < n; i += gridDim.x * kBuildBlock) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (k < dims) {
                float p = __uint_as_float(cols[(2u + k) * cap + i]);
                float q = is_point ? p : __uint_as_float(cols[(6u + k) * cap + i]);
                lo[k] = fminf(lo[k], p);
                hi[k] = fmaxf(hi[k], q);
                if (!is_point)
                    ext[k] = fmaxf(ext[k], q - p);
            }
        }
    }
    __shared__ float s[kBuildBlock / 64][kBoxPartial];
    const unsigned lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float a = wave_min(lo[k]), b = wave_max(hi[k]), e = wave_max(ext[k]);
        if (lane == 0) {
            s[w][k] = a;
            s[w][3 + k] = b;
            s[w][6 + k] = e;
        }
    }
    __syncthreads();
    if (threadIdx.x < kBoxPartial) {
        const unsigned k = threadIdx.x;
        float v = s[0][k];
        for (unsigned j = 1; j < kBuildBlock / 64; ++j)
            v = k < 3 ? fminf(v, s[j][k]) : fmaxf(v, s[j][k]);
        partials[blockIdx.x * kBoxPartial + k] = v;
    }
}

/* Grid of one box.  Per used axis
 *   size  = (max - min) + part * 0.001f
 *   count = (int)ceilf(size / part), at least 1
 * coarsened by the smallest integer k with every ceil(count / k) within the
 * Morton range and their product within max_cells; then
 *   cell_size = size / (float)count. */
__device__ void grid_geometry(const float *lo, const float *hi, const float *part, int dims,
                              unsigned long long max_cells, int order_bits, TayGridGeom *__restrict__ geom) {
    float size[3] = {1.0f, 1.0f, 1.0f};
    unsigned long long count[3] = {1, 1, 1};
    for (int k = 0; k < dims; ++k) {
        size[k] = (hi[k] - lo[k]) + part[k] * 0.001f;
        int c = (int)ceilf(size[k] / part[k]);
        count[k] = c < 1 ? 1ull : (unsigned long long)c;
    }
    const unsigned long long limit = dims == 3 ? 1024ull : dims == 2 ? 65536ull : 0xFFFFFFFFull;
    /* fits(k) is monotone in k, so the search may start at any lower bound */
    unsigned long long k = 1;
    double product = 1.0;
    for (int a = 0; a < dims; ++a) {
        unsigned long long need = (count[a] + limit - 1) / limit;
        k = need > k ? need : k;
        product *= (double)count[a];
    }
    if (product > (double)max_cells) {
        /* far beyond the cap: start the search near the answer */
        double root = pow(product / (double)(max_cells ? max_cells : 1ull), 1.0 / (double)dims);
        if (root > 2.0 && (unsigned long long)root - 1ull > k)
            k = (unsigned long long)root - 1ull;
    }
    for (;; ++k) {
        unsigned long long total = 1;
        bool fits = true;
        for (int a = 0; a < dims; ++a) {
            unsigned long long c = (count[a] + k - 1) / k;
            fits = fits && c <= limit;
            total *= c;
        }
        if (fits && total <= max_cells)
            break;
    }
    unsigned long long cells = 1;
    for (int a = 0; a < 4; ++a) {
        if (a < dims) {
            unsigned long long c = (count[a] + k - 1) / k;
            geom->count[a] = (int)c;
            geom->cell_size[a] = size[a] / (float)(int)c;
            geom->origin[a] = lo[a];
            geom->box_min[a] = lo[a];
            geom->box_max[a] = hi[a];
            cells *= c;
        } else {
            geom->count[a] = 1;
            geom->cell_size[a] = 1.0f;
            geom->origin[a] = 0.0f;
            geom->box_min[a] = 0.0f;
            geom->box_max[a] = 0.0f;
        }
        geom->max_extent[a] = 0.0f;
    }
    geom->cells = cells;
    /* the finest sub-cell order whose key still fits the Morton range */
    int s = order_bits;
    for (; s > 0; --s) {
        bool fits = true;
        for (int a = 0; a < dims; ++a)
            fits = fits && ((unsigned long long)geom->count[a] << s) <= limit;
        if (fits)
            break;
    }
    geom->order_bits = s;
}

/* One wavefront: reduce the block boxes and derive the grid; the largest
 * extent of non-point agents goes into the geometry. */
__global__ __launch_bounds__(64) void k_geometry(const float *__restrict__ partials, unsigned nparts, int dims,
                                                 float px, float py, float pz, unsigned long long max_cells,
                                                 int order_bits, int box_fixed, float fx0, float fy0, float fz0,
                                                 float fx1, float fy1, float fz1, TayGridGeom *__restrict__ geom) {
    const unsigned lane = threadIdx.x;
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    float ext[3] = {0.0f, 0.0f, 0.0f};
    for (unsigned b = lane; b < nparts; b += 64) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(lo[k], partials[b * kBoxPartial + k]);
            hi[k] = fmaxf(hi[k], partials[b * kBoxPartial + 3 + k]);
            ext[k] = fmaxf(ext[k], partials[b * kBoxPartial + 6 + k]);
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = wave_min(lo[k]);
        hi[k] = wave_max(hi[k]);
        ext[k] = wave_max(ext[k]);
    }
    if (lane != 0)
        return;
    if (box_fixed) {
        lo[0] = fx0; lo[1] = fy0; lo[2] = fz0;
        hi[0] = fx1; hi[1] = fy1; hi[2] = fz1;
    }
    const float part[3] = {px, py, pz};
    grid_geometry(lo, hi, part, dims, max_cells, order_bits, geom);
    for (int k = 0; k < dims; ++k)
        geom->max_extent[k] = ext[k];
}

__device__ __forceinline__ int cell_axis(float p, float origin, float size, int count) {
    int c = (int)floorf((p - origin) / size);
    c = c < 0 ? 0 : c;
    return c > count - 1 ? count - 1 : c;
}

/* Storage-order key of agent i: per axis the cell index c = (int)floorf(q)
 * clamped, q = (p - origin) / size as in tay/cpu_grid.c:18-41, refined by
 * s = order_bits sub-cell bits, sub = (int)floorf((q - c) * 2^s) clamped to
 * [0, 2^s - 1]; the key is the Morton interleave of (c << s) | sub.  With
 * s = 0 it is the cell's Morton key (tay/cpu_z_grid.c:37-83). */
__device__ __forceinline__ unsigned storage_key(const unsigned *__restrict__ cols, unsigned cap, unsigned i,
                                                int dims, const TayGridGeom *__restrict__ g) {
    const int s = g->order_bits;
    int idx[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (k < dims) {
            const float q = (__uint_as_float(cols[(2u + k) * cap + i]) - g->origin[k]) / g->cell_size[k];
            int c = (int)floorf(q);
            c = c < 0 ? 0 : c;
            c = c > g->count[k] - 1 ? g->count[k] - 1 : c;
            int sub = 0;
            if (s) {
                sub = (int)floorf((q - (float)c) * (float)(1 << s));
                sub = sub < 0 ? 0 : sub;
                sub = sub > (1 << s) - 1 ? (1 << s) - 1 : sub;
            }
            idx[k] = (c << s) | sub;
        }
    return morton(idx[0], idx[1], idx[2], dims);
}

__global__ __launch_bounds__(kBuildBlock) void k_keys(const unsigned *__restrict__ cols, unsigned cap, unsigned n,
                                                      int dims, const TayGridGeom *__restrict__ g,
                                                      unsigned *__restrict__ keys) {
    const unsigned i = blockIdx.x * kBuildBlock + threadIdx.x;
    if (i >= n)
        return;
    keys[i] = storage_key(cols, cap, i, dims, g);
}

/* k_keys and k_ranges_clear in one launch: slot keys for i < n, and the
 * step's cell table cleared by a grid-stride loop (nothing reads the table
 * between here and k_ranges). */
__global__ __launch_bounds__(kBuildBlock) void k_keys_clear(const unsigned *__restrict__ cols, unsigned cap,
                                                            unsigned n, int dims, const TayGridGeom *__restrict__ g,
                                                            unsigned *__restrict__ keys,
                                                            TayCellRange *__restrict__ ranges) {
    const unsigned i = blockIdx.x * kBuildBlock + threadIdx.x;
    if (i < n)
        keys[i] = storage_key(cols, cap, i, dims, g);
    const unsigned long long cells = g->cells;
    for (unsigned long long c = i; c < cells; c += (unsigned long long)gridDim.x * kBuildBlock) {
        ranges[c].beg = 0;
        ranges[c].end = 0;
    }
}

/* One radix tile of a segmented sort: keys [start, start + count) never
 * leave their segment; the tile's count of digit d goes to histogram word
 * hist_base + d * hist_stride.  Laid out segment-major (for each segment,
 * its digits, for each digit its tiles), one exclusive scan of the
 * histogram then gives every key its destination inside its own segment. */
struct TaySortTile {
    unsigned start, count, hist_base, hist_stride;
};

/* Tile geometry of block b: a segment tile, or tile b of one unsegmented
 * sort of n keys over nblocks tiles (histogram digit-major:
 * hist[d * nblocks + b]). */
__device__ __forceinline__ TaySortTile sort_tile(const TaySortTile *__restrict__ seg, unsigned b, unsigned n,
                                                 unsigned nblocks) {
    if (seg)
        return seg[b];
    TaySortTile t;
    t.start = b * kRadixTile;
    t.count = n - t.start < kRadixTile ? n - t.start : kRadixTile;
    t.hist_base = b;
    t.hist_stride = nblocks;
    return t;
}

/* Digit histogram of one tile of at most 4096 keys per block. */
/* The radix pass at `shift` leaves the order as it is when every real key
 * is below 2^shift - 1: their digits from `shift` up are zero, and a
 * padding key (all ones, a batch's empty slots) already sorts after them on
 * its lower bits.  key_max is the largest real key of the sort (nullptr:
 * no pass is skipped); the first pass always runs. */
__device__ __forceinline__ bool radix_pass_idle(const unsigned *key_max, unsigned shift) {
    return key_max && shift && *key_max < (1u << shift) - 1u;
}

__global__ __launch_bounds__(kRadixBlock) void k_radix_hist(const unsigned *__restrict__ keys, unsigned n,
                                                            unsigned shift, unsigned *__restrict__ hist,
                                                            unsigned nblocks, const TaySortTile *__restrict__ seg,
                                                            const unsigned *__restrict__ key_max) {
    if (radix_pass_idle(key_max, shift))
        return;
    __shared__ unsigned h[kRadix];
    if (threadIdx.x < kRadix)
        h[threadIdx.x] = 0;
    __syncthreads();
    const TaySortTile tile = sort_tile(seg, blockIdx.x, n, nblocks);
    for (unsigned r = threadIdx.x; r < tile.count; r += kRadixBlock)
        atomicAdd(&h[(keys[tile.start + r] >> shift) & (kRadix - 1)], 1u);
    __syncthreads();
    if (threadIdx.x < kRadix)
        hist[tile.hist_base + threadIdx.x * tile.hist_stride] = h[threadIdx.x];
}

/* Exclusive scan of the digit-major histogram in place, in two steps over
 * kScanBlocks contiguous segments: segment sums, then each segment scanned
 * in coalesced rounds from its carry-in (the sum of the segment sums before
 * it, formed by each block). */
__device__ __forceinline__ unsigned wave_inclusive_sum(unsigned v, unsigned lane) {
    for (unsigned o = 1; o < 64; o <<= 1) {
        unsigned u = __shfl_up(v, o);
        v += lane >= o ? u : 0u;
    }
    return v;
}

__device__ __forceinline__ unsigned segment_begin(unsigned s, unsigned total, unsigned nseg) {
    return (unsigned)(((unsigned long long)total * s) / nseg);
}

__global__ __launch_bounds__(kBuildBlock) void k_scan_reduce(const unsigned *__restrict__ hist, unsigned total,
                                                             unsigned *__restrict__ sums,
                                                             const unsigned *__restrict__ key_max, unsigned shift) {
    if (radix_pass_idle(key_max, shift))
        return;
    __shared__ unsigned s[kBuildBlock / 64];
    const unsigned beg = segment_begin(blockIdx.x, total, gridDim.x);
    const unsigned end = segment_begin(blockIdx.x + 1, total, gridDim.x);
    unsigned v = 0;
    for (unsigned i = beg + threadIdx.x; i < end; i += kBuildBlock)
        v += hist[i];
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o);
    if ((threadIdx.x & 63u) == 0)
        s[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned t = 0;
        for (unsigned w = 0; w < kBuildBlock / 64; ++w)
            t += s[w];
        sums[blockIdx.x] = t;
    }
}


__global__ __launch_bounds__(kBuildBlock) void k_scan_down(unsigned *__restrict__ hist, unsigned total,
                                                           const unsigned *__restrict__ sums,
                                                           const unsigned *__restrict__ key_max, unsigned shift) {
    if (radix_pass_idle(key_max, shift))
        return;
    __shared__ unsigned s[kBuildBlock / 64];
    const unsigned lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const unsigned beg = segment_begin(blockIdx.x, total, gridDim.x);
    const unsigned end = segment_begin(blockIdx.x + 1, total, gridDim.x);
    /* carry-in: the sums of the segments before this one (at most 64, one
     * per lane), added up by every wavefront */
    unsigned carry = wave_inclusive_sum(lane < blockIdx.x ? sums[lane] : 0u, lane);
    carry = __shfl(carry, 63);
    for (unsigned base = beg; base < end; base += kBuildBlock) {
        const unsigned i = base + threadIdx.x;
        const unsigned v = i < end ? hist[i] : 0u;
        const unsigned inc = wave_inclusive_sum(v, lane);
        if (lane == 63u)
            s[w] = inc;
        __syncthreads();
        unsigned before = 0, round = 0;
        for (unsigned k = 0; k < kBuildBlock / 64; ++k) {
            before += k < w ? s[k] : 0u;
            round += s[k];
        }
        if (i < end)
            hist[i] = carry + before + inc - v;
        carry += round;
        __syncthreads();
    }
}

/* Stable scatter of one tile, through LDS.  Wave w of the 16 owns the
 * contiguous sixteenth [w * 256, (w + 1) * 256) of the tile; inside a wave,
 * 4 rounds of 64 keys go in order and lanes with equal digits are ranked by
 * lane (ballot match).  Sixteen wavefronts per tile keep enough loads in
 * flight when the tiles are few (a 1M-key sort has 245).  Every key first goes to its place in the tile sorted by
 * digit (in LDS); then the block writes the sorted tile out in order, so
 * the keys of one digit leave as one contiguous run of the output instead
 * of one scattered word each. */
__global__ __launch_bounds__(kRadixBlock) void k_radix_scatter(const unsigned *__restrict__ keys_in,
                                                               const unsigned *__restrict__ vals_in,
                                                               unsigned *__restrict__ keys_out,
                                                               unsigned *__restrict__ vals_out, unsigned n,
                                                               unsigned shift, const unsigned *__restrict__ offsets,
                                                               unsigned nblocks, int identity_vals,
                                                               const TaySortTile *__restrict__ seg,
                                                               const unsigned *__restrict__ key_max) {
    __shared__ unsigned wave_hist[kRadixWaves][kRadix];
    __shared__ unsigned lbase[kRadixWaves][kRadix];   /* next tile slot of (wave, digit) */
    __shared__ unsigned delta[kRadix];                /* output slot minus tile slot of a digit */
    __shared__ unsigned wave_sum[kRadix / 64];
    __shared__ unsigned lkeys[kRadixTile], lvals[kRadixTile];
    const unsigned t = threadIdx.x, lane = t & 63u, w = t >> 6;
    if (radix_pass_idle(key_max, shift)) {
        /* an idle pass still moves the tile to the other buffer */
        const TaySortTile tile = sort_tile(seg, blockIdx.x, n, nblocks);
        for (unsigned r = t; r < tile.count; r += kRadixBlock) {
            keys_out[tile.start + r] = keys_in[tile.start + r];
            vals_out[tile.start + r] = vals_in[tile.start + r];
        }
        return;
    }
    for (unsigned k = t; k < kRadixWaves * kRadix; k += kRadixBlock)
        (&wave_hist[0][0])[k] = 0;
    __syncthreads();
    const TaySortTile tile = sort_tile(seg, blockIdx.x, n, nblocks);
    const unsigned end = tile.start + tile.count;
    const unsigned sub = tile.start + w * (kRadixRounds * 64);
    for (unsigned r = 0; r < kRadixRounds; ++r) {
        const unsigned idx = sub + r * 64 + lane;
        if (idx < end)
            atomicAdd(&wave_hist[w][(keys_in[idx] >> shift) & (kRadix - 1)], 1u);
    }
    __syncthreads();
    /* threads t < kRadix: thread t is digit t, with its count in the tile,
     * the tile slots before it (scan over the digits), then per wave */
    unsigned c = 0, inc = 0;
    if (t < kRadix) {
        for (unsigned ww = 0; ww < kRadixWaves; ++ww)
            c += wave_hist[ww][t];
        inc = wave_inclusive_sum(c, lane);
        if (lane == 63u)
            wave_sum[w] = inc;
    }
    __syncthreads();
    if (t < kRadix) {
        unsigned start = inc - c;
        for (unsigned ww = 0; ww < w; ++ww)
            start += wave_sum[ww];
        delta[t] = offsets[tile.hist_base + t * tile.hist_stride] - start;
        for (unsigned ww = 0; ww < kRadixWaves; ++ww) {
            lbase[ww][t] = start;
            start += wave_hist[ww][t];
        }
    }
    __syncthreads();
    const unsigned long long below = (1ull << lane) - 1ull;
    for (unsigned r = 0; r < kRadixRounds; ++r) {
        const unsigned idx = sub + r * 64 + lane;
        const bool valid = idx < end;
        const unsigned key = valid ? keys_in[idx] : 0u;
        const unsigned val = valid ? (identity_vals ? idx : vals_in[idx]) : 0u;
        const unsigned d = (key >> shift) & (kRadix - 1);
        unsigned long long m = __ballot(valid);
#pragma unroll
        for (unsigned b = 0; b < kRadixBits; ++b) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long vote = __ballot(bit);
            m &= bit ? vote : ~vote;
        }
        const unsigned rank = __popcll(m & below);
        if (valid) {
            const unsigned pos = lbase[w][d] + rank;
            lkeys[pos] = key;
            lvals[pos] = val;
        }
        __builtin_amdgcn_wave_barrier();
        if (valid && rank == 0)
            lbase[w][d] += __popcll(m);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    for (unsigned p = t; p < tile.count; p += kRadixBlock) {
        const unsigned key = lkeys[p], val = lvals[p];
        const unsigned pos = delta[(key >> shift) & (kRadix - 1)] + p;
        keys_out[pos] = key;
        vals_out[pos] = val;
    }
}

/* Stable LSD radix sort of at most kSmallSort keys by one block, every pass
 * in LDS, kSmallBits (9) per pass: one launch instead of four histogram,
 * scan and scatter rounds, for the small states whose step is a chain of
 * short launches (keys below 2^18, e.g. a 16^3-cell grid with two sub-cell
 * bits, take two passes).
 * Wave w owns the contiguous slice [w * per, (w + 1) * per) of the current
 * order; per pass each lane keeps its keys in registers, ranks them inside
 * the wave round by round (ballot match, as k_radix_scatter), the digit
 * starts are scanned over (digit, wave), and the keys are written back to
 * the LDS arrays at their new places; passes above the largest key's top
 * digit are skipped.  The source index of slot i starts as i.  Sorts
 * keys[0, n) in place and writes the source indices to vals. */

__global__ __launch_bounds__(kRadixBlock) void k_small_sort(unsigned *__restrict__ keys,
                                                            unsigned *__restrict__ vals, unsigned n, int dims,
                                                            const TayGridGeom *__restrict__ g,
                                                            TayCellRange *__restrict__ ranges) {
    __shared__ unsigned lkeys[kSmallSort], lvals[kSmallSort];
    __shared__ unsigned cnt[kRadixWaves][kSmallRadix];
    __shared__ unsigned wave_sum[kSmallRadix / 64];
    __shared__ unsigned wave_max_key[kRadixWaves];
    const unsigned t = threadIdx.x, lane = t & 63u, w = t >> 6;
    unsigned kmax = 0;
    {
        /* every load issued before the first is used: one memory latency */
        unsigned k[kSmallPerLane];
#pragma unroll
        for (unsigned r = 0; r < kSmallPerLane; ++r) {
            const unsigned i = t + r * kRadixBlock;
            k[r] = i < n ? keys[i] : 0u;
        }
#pragma unroll
        for (unsigned r = 0; r < kSmallPerLane; ++r) {
            const unsigned i = t + r * kRadixBlock;
            if (i < n) {
                lkeys[i] = k[r];
                lvals[i] = i;
                kmax = k[r] > kmax ? k[r] : kmax;
            }
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned u = __shfl_xor(kmax, o);
        kmax = u > kmax ? u : kmax;
    }
    if (lane == 0u)
        wave_max_key[w] = kmax;
    __syncthreads();
    kmax = 0;
    for (unsigned ww = 0; ww < kRadixWaves; ++ww)
        kmax = wave_max_key[ww] > kmax ? wave_max_key[ww] : kmax;
    const unsigned rounds = (n + kRadixWaves * 64u - 1u) / (kRadixWaves * 64u);   /* <= kSmallPerLane */
    const unsigned per = rounds * 64u;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (unsigned pass = 0; pass * kSmallBits < 32u; ++pass) {
        const unsigned shift = pass * kSmallBits;
        /* every digit from here up is zero: the remaining passes would keep
         * the order (stable), so the keys are sorted */
        if (pass > 0u && (kmax >> shift) == 0u)
            break;
        for (unsigned k = t; k < kRadixWaves * kSmallRadix; k += kRadixBlock)
            (&cnt[0][0])[k] = 0;
        __syncthreads();
        unsigned key[kSmallPerLane], val[kSmallPerLane], pos[kSmallPerLane];
#pragma unroll
        for (unsigned r = 0; r < kSmallPerLane; ++r) {
            key[r] = val[r] = pos[r] = 0u;
            /* rounds is the same for the whole block: the rounds past it
             * are skipped, not ranked with every lane masked off */
            if (r >= rounds)
                continue;
            const unsigned idx = w * per + r * 64u + lane;
            const bool valid = idx < n;
            key[r] = valid ? lkeys[idx] : 0u;
            val[r] = valid ? lvals[idx] : 0u;
            const unsigned d = (key[r] >> shift) & (kSmallRadix - 1);
            unsigned long long m = __ballot(valid);
#pragma unroll
            for (unsigned b = 0; b < kSmallBits; ++b) {
                const bool bit = (d >> b) & 1u;
                const unsigned long long vote = __ballot(bit);
                m &= bit ? vote : ~vote;
            }
            /* slot inside the wave's run of digit d, then the run grows */
            pos[r] = valid ? cnt[w][d] + (unsigned)__popcll(m & below) : 0u;
            __builtin_amdgcn_wave_barrier();
            if (valid && (m & below) == 0ull)
                cnt[w][d] += (unsigned)__popcll(m);
            __builtin_amdgcn_wave_barrier();
        }
        __syncthreads();
        /* threads t < kSmallRadix: digit t's start, then each wave's run start */
        unsigned c = 0, inc = 0;
        if (t < kSmallRadix) {
            for (unsigned ww = 0; ww < kRadixWaves; ++ww)
                c += cnt[ww][t];
            inc = wave_inclusive_sum(c, lane);
            if (lane == 63u)
                wave_sum[w] = inc;
        }
        __syncthreads();
        if (t < kSmallRadix) {
            unsigned start = inc - c;
            for (unsigned ww = 0; ww < w; ++ww)
                start += wave_sum[ww];
            for (unsigned ww = 0; ww < kRadixWaves; ++ww) {
                const unsigned run = cnt[ww][t];
                cnt[ww][t] = start;
                start += run;
            }
        }
        __syncthreads();
#pragma unroll
        for (unsigned r = 0; r < kSmallPerLane; ++r) {
            const unsigned idx = w * per + r * 64u + lane;
            if (r < rounds && idx < n) {
                const unsigned dst = cnt[w][(key[r] >> shift) & (kSmallRadix - 1)] + pos[r];
                lkeys[dst] = key[r];
                lvals[dst] = val[r];
            }
        }
        __syncthreads();
    }
    /* the sorted keys and source indices, and (as k_ranges does from them)
     * the [beg, end) of every occupied cell, into the table k_keys_clear
     * cleared */
    for (unsigned i = t; i < n; i += kRadixBlock) {
        keys[i] = lkeys[i];
        vals[i] = lvals[i];
        const unsigned key = cell_of_key(lkeys[i], dims, g);
        const bool first = i == 0 || cell_of_key(lkeys[i - 1], dims, g) != key;
        const bool last = i + 1 == n || cell_of_key(lkeys[i + 1], dims, g) != key;
        if (first || last) {
            TayCellRange *r = ranges + morton_to_cell(key, dims, g);
            if (first)
                r->beg = i;
            if (last)
                r->end = i + 1;
        }
    }
}

__global__ __launch_bounds__(kBuildBlock) void k_ranges_clear(TayCellRange *__restrict__ ranges,
                                                              const TayGridGeom *__restrict__ g) {
    const unsigned long long cells = g->cells;
    for (unsigned long long c = blockIdx.x * (unsigned long long)kBuildBlock + threadIdx.x; c < cells;
         c += (unsigned long long)gridDim.x * kBuildBlock) {
        ranges[c].beg = 0;
        ranges[c].end = 0;
    }
}

__global__ __launch_bounds__(kBuildBlock) void k_ranges(const unsigned *__restrict__ keys, unsigned n, int dims,
                                                        const TayGridGeom *__restrict__ g,
                                                        TayCellRange *__restrict__ ranges) {
    const unsigned i = blockIdx.x * kBuildBlock + threadIdx.x;
    if (i >= n)
        return;
    const unsigned key = cell_of_key(keys[i], dims, g);
    const bool first = i == 0 || cell_of_key(keys[i - 1], dims, g) != key;
    const bool last = i + 1 == n || cell_of_key(keys[i + 1], dims, g) != key;
    if (first || last) {
        TayCellRange *r = ranges + morton_to_cell(key, dims, g);
        if (first)
            r->beg = i;
        if (last)
            r->end = i + 1;
    }
}

/* dst column words [w0, words) of slot i from slot `from` of src, eight
 * loads in flight before their stores (a runtime-length loop would wait on
 * every load before the next). */
__device__ __forceinline__ void gather_words(const unsigned *__restrict__ src, unsigned *__restrict__ dst,
                                             unsigned cap, unsigned from, unsigned i, unsigned w0, unsigned words) {
    unsigned w = w0;
    for (; w + 8u <= words; w += 8u) {
        unsigned t[8];
#pragma unroll
        for (unsigned k = 0; k < 8u; ++k)
            t[k] = src[(w + k) * cap + from];
#pragma unroll
        for (unsigned k = 0; k < 8u; ++k)
            dst[(w + k) * cap + i] = t[k];
    }
    if (w + 4u <= words) {
        unsigned t[4];
#pragma unroll
        for (unsigned k = 0; k < 4u; ++k)
            t[k] = src[(w + k) * cap + from];
#pragma unroll
        for (unsigned k = 0; k < 4u; ++k)
            dst[(w + k) * cap + i] = t[k];
        w += 4u;
    }
    for (; w < words; ++w)
        dst[w * cap + i] = src[w * cap + from];
}

/* Gather every column into key order; word 0 becomes the key (part_i) and
 * word 1 the rank inside the cell (cell_agent_i). */
__global__ __launch_bounds__(kBuildBlock) void k_permute(const unsigned *__restrict__ keys,
                                                         const unsigned *__restrict__ vals, unsigned n,
                                                         unsigned words, unsigned cap, int dims,
                                                         const TayGridGeom *__restrict__ g,
                                                         const TayCellRange *__restrict__ ranges,
                                                         const unsigned *__restrict__ src,
                                                         unsigned *__restrict__ dst) {
    const unsigned i = blockIdx.x * kBuildBlock + threadIdx.x;
    if (i >= n)
        return;
    const unsigned key = keys[i];
    const unsigned from = vals[i];
    dst[i] = cell_of_key(key, dims, g);
    dst[cap + i] = i - ranges[key_to_cell(key, dims, g)].beg;
    gather_words(src, dst, cap, from, i, 2u, words);
}

__global__ __launch_bounds__(kBuildBlock) void k_aos_to_soa(const unsigned *__restrict__ aos,
                                                            unsigned *__restrict__ cols, unsigned n, unsigned words,
                                                            unsigned cap) {
    const unsigned long long total = (unsigned long long)n * words;
    for (unsigned long long e = blockIdx.x * (unsigned long long)kBuildBlock + threadIdx.x; e < total;
         e += (unsigned long long)gridDim.x * kBuildBlock) {
        const unsigned agent = (unsigned)(e / words), w = (unsigned)(e % words);
        cols[(unsigned long long)w * cap + agent] = aos[e];
    }
}

__global__ __launch_bounds__(kBuildBlock) void k_soa_to_aos(const unsigned *__restrict__ cols,
                                                            unsigned *__restrict__ aos, unsigned n, unsigned words,
                                                            unsigned cap) {
    const unsigned long long total = (unsigned long long)n * words;
    for (unsigned long long e = blockIdx.x * (unsigned long long)kBuildBlock + threadIdx.x; e < total;
         e += (unsigned long long)gridDim.x * kBuildBlock) {
        const unsigned agent = (unsigned)(e / words), w = (unsigned)(e % words);
        aos[e] = cols[(unsigned long long)w * cap + agent];
    }
}

/* Streaming copy of n 16-byte vectors: the calibration kernel for HBM
 * counters and a measured copy bandwidth. */
__global__ __launch_bounds__(kBuildBlock) void k_copy_probe(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                            unsigned long long n) {
    for (unsigned long long i = blockIdx.x * (unsigned long long)kBuildBlock + threadIdx.x; i < n;
         i += (unsigned long long)gridDim.x * kBuildBlock)
        dst[i] = src[i];
}

unsigned blocks_for(unsigned long long work, unsigned per_block, unsigned cap_blocks) {
    unsigned long long b = (work + per_block - 1) / per_block;
    if (b == 0)
        b = 1;
    return (unsigned)(b < cap_blocks ? b : cap_blocks);
}

bool launch_ok(const char *what) {
    return tay::hip_ok(hipGetLastError(), what);
}

}  // namespace

#endif
