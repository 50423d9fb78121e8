/*
 * gpu_space.hip - device storage of agent groups and the ahead-of-time Z-grid
 * build for gfx950.
 *
 * Agents live as structure-of-arrays columns of 32-bit words.  A sort
 * (tay_gpu_space_sort) rebuilds the grid from the current positions without
 * any host synchronisation:
 *
 *   box_partial  per-block min/max of the positions      (tay/space.c:30-48)
 *   geometry     one wavefront: box, cell counts/sizes    (tay/ocl_grid.c:128-136)
 *   keys         per agent: cell indices, Morton key      (tay/cpu_grid.c:18-41,
 *                                                          tay/cpu_z_grid.c:37-83)
 *   radix x4     stable LSD radix sort of the keys, 8 bits per pass
 *   ranges       [beg, end) of every occupied cell, from key boundaries
 *   permute      gather every column into key order, write the tag
 *
 * The stable sort keeps agents of a cell in their previous storage order, as
 * the reference's counting sorts do (tay/cpu_grid.c:154-188,
 * tay/cpu_z_grid.c:192-215); the reference's OpenCL grid ranks agents with
 * atomic_inc instead (tay/ocl_grid.c:147), which is not deterministic.
 *
 * Every floating-point step of the geometry and key kernels is the plain IEEE
 * sequence of the reference (the library is built with -ffp-contract=off), so
 * cell indices and keys match the CPU statement bit for bit.
 */
#include "internal.h"

#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace tay {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...) {
    char buf[2048];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

bool hip_ok(hipError_t e, const char *what) {
    if (e == hipSuccess)
        return true;
    set_error("%s: %s", what, hipGetErrorString(e));
    return false;
}

static hipStream_t g_stream = nullptr;
static int g_device_state = 0;  /* 0 unknown, 1 ready, -1 unavailable */

hipStream_t stream() {
    if (g_device_state == 1)
        return g_stream;
    if (g_device_state == -1) {
        set_error("no usable GPU");
        return nullptr;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        g_device_state = -1;
        set_error("no usable GPU");
        return nullptr;
    }
    if (!hip_ok(hipSetDevice(0), "hipSetDevice") ||
        !hip_ok(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking), "hipStreamCreate")) {
        g_device_state = -1;
        return nullptr;
    }
    g_device_state = 1;
    return g_stream;
}

bool device_alloc(void **ptr, size_t bytes, const char *what) {
    *ptr = nullptr;
    if (!stream())
        return false;
    if (bytes == 0)
        bytes = 4;
    return hip_ok(hipMalloc(ptr, bytes), what);
}

void device_free(void *ptr) {
    if (ptr)
        (void)hipFree(ptr);
}

}  // namespace tay

using namespace tay;

/* ------------------------------------------------------------------ kernels */

namespace {

constexpr unsigned kBuildBlock = 256;
constexpr unsigned kRadixBits = 8;
constexpr unsigned kRadix = 1u << kRadixBits;
constexpr unsigned kRadixPasses = 4;            /* 32-bit keys */
constexpr unsigned kRadixWaves = kBuildBlock / 64;
constexpr unsigned kRadixRounds = 16;
constexpr unsigned kRadixTile = kRadixWaves * kRadixRounds * 64;   /* keys per block */
constexpr unsigned kScanBlocks = 64;             /* segments of the histogram scan, at most one wavefront's lanes */

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

/* Row-major cell id (x fastest) of a Morton key. */
__device__ __forceinline__ unsigned long long key_to_cell(unsigned key, int dims, const TayGridGeom *g) {
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
 * non-point agents the max corner at words 6..8). */
__global__ __launch_bounds__(kBuildBlock) void k_box_partial(const unsigned *__restrict__ cols, unsigned cap,
                                                             unsigned n, int dims, int is_point,
                                                             float *__restrict__ partials) {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (unsigned i = blockIdx.x * kBuildBlock + threadIdx.x; i < n; i += gridDim.x * kBuildBlock) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (k < dims) {
                float p = __uint_as_float(cols[(2u + k) * cap + i]);
                float q = is_point ? p : __uint_as_float(cols[(6u + k) * cap + i]);
                lo[k] = fminf(lo[k], p);
                hi[k] = fmaxf(hi[k], q);
            }
        }
    }
    __shared__ float s[kBuildBlock / 64][6];
    const unsigned lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float a = wave_min(lo[k]), b = wave_max(hi[k]);
        if (lane == 0) {
            s[w][k] = a;
            s[w][3 + k] = b;
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const unsigned k = threadIdx.x;
        float v = s[0][k];
        for (unsigned j = 1; j < kBuildBlock / 64; ++j)
            v = k < 3 ? fminf(v, s[j][k]) : fmaxf(v, s[j][k]);
        partials[blockIdx.x * 6u + k] = v;
    }
}

/* One wavefront: reduce the block boxes and derive the grid.  Per used axis
 *   size  = (max - min) + part * 0.001f
 *   count = (int)ceilf(size / part), at least 1
 * coarsened by the smallest integer k with every ceil(count / k) within the
 * Morton range and their product within max_cells; then
 *   cell_size = size / (float)count. */
__global__ __launch_bounds__(64) void k_geometry(const float *__restrict__ partials, unsigned nparts, int dims,
                                                 float px, float py, float pz, unsigned long long max_cells,
                                                 int box_fixed, float fx0, float fy0, float fz0, float fx1,
                                                 float fy1, float fz1, TayGridGeom *__restrict__ geom) {
    const unsigned lane = threadIdx.x;
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (unsigned b = lane; b < nparts; b += 64) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(lo[k], partials[b * 6u + k]);
            hi[k] = fmaxf(hi[k], partials[b * 6u + 3 + k]);
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = wave_min(lo[k]);
        hi[k] = wave_max(hi[k]);
    }
    if (lane != 0)
        return;
    if (box_fixed) {
        lo[0] = fx0; lo[1] = fy0; lo[2] = fz0;
        hi[0] = fx1; hi[1] = fy1; hi[2] = fz1;
    }
    const float part[3] = {px, py, pz};
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
    double root = pow(product / (double)(max_cells ? max_cells : 1ull), 1.0 / (double)dims);
    if (root > 2.0 && (unsigned long long)root - 1ull > k)
        k = (unsigned long long)root - 1ull;
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
    }
    geom->cells = cells;
}

__device__ __forceinline__ int cell_axis(float p, float origin, float size, int count) {
    int c = (int)floorf((p - origin) / size);
    c = c < 0 ? 0 : c;
    return c > count - 1 ? count - 1 : c;
}

__global__ __launch_bounds__(kBuildBlock) void k_keys(const unsigned *__restrict__ cols, unsigned cap, unsigned n,
                                                      int dims, const TayGridGeom *__restrict__ g,
                                                      unsigned *__restrict__ keys) {
    const unsigned i = blockIdx.x * kBuildBlock + threadIdx.x;
    if (i >= n)
        return;
    int idx[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (k < dims)
            idx[k] = cell_axis(__uint_as_float(cols[(2u + k) * cap + i]), g->origin[k], g->cell_size[k], g->count[k]);
    keys[i] = morton(idx[0], idx[1], idx[2], dims);
}

/* Digit histogram of one 4096-key tile per block, stored digit-major:
 * hist[d * nblocks + block]. */
__global__ __launch_bounds__(kBuildBlock) void k_radix_hist(const unsigned *__restrict__ keys, unsigned n,
                                                            unsigned shift, unsigned *__restrict__ hist,
                                                            unsigned nblocks) {
    __shared__ unsigned h[kRadix];
    h[threadIdx.x] = 0;
    __syncthreads();
    const unsigned base = blockIdx.x * kRadixTile;
    for (unsigned r = 0; r < kRadixTile / kBuildBlock; ++r) {
        const unsigned idx = base + r * kBuildBlock + threadIdx.x;
        if (idx < n)
            atomicAdd(&h[(keys[idx] >> shift) & (kRadix - 1)], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

/* Exclusive scan of the digit-major histogram in place, in three steps over
 * kScanBlocks contiguous segments: segment sums, a one-wavefront scan of the
 * sums, then each segment scanned in coalesced rounds from its carry-in. */
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
                                                             unsigned *__restrict__ sums) {
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

__global__ __launch_bounds__(64) void k_scan_top(unsigned *__restrict__ sums, unsigned n) {
    const unsigned lane = threadIdx.x;
    const unsigned v = lane < n ? sums[lane] : 0u;
    const unsigned inc = wave_inclusive_sum(v, lane);
    if (lane < n)
        sums[lane] = inc - v;
}

__global__ __launch_bounds__(kBuildBlock) void k_scan_down(unsigned *__restrict__ hist, unsigned total,
                                                           const unsigned *__restrict__ sums) {
    __shared__ unsigned s[kBuildBlock / 64];
    const unsigned lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const unsigned beg = segment_begin(blockIdx.x, total, gridDim.x);
    const unsigned end = segment_begin(blockIdx.x + 1, total, gridDim.x);
    unsigned carry = sums[blockIdx.x];
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

/* Stable scatter of one tile.  Wave w owns the contiguous quarter
 * [w * 1024, (w + 1) * 1024) of the tile and places it after the same
 * digits of the waves before it; inside a wave, 16 rounds of 64 keys go in
 * order and lanes with equal digits are ranked by lane (ballot match). */
__global__ __launch_bounds__(kBuildBlock) void k_radix_scatter(const unsigned *__restrict__ keys_in,
                                                               const unsigned *__restrict__ vals_in,
                                                               unsigned *__restrict__ keys_out,
                                                               unsigned *__restrict__ vals_out, unsigned n,
                                                               unsigned shift, const unsigned *__restrict__ offsets,
                                                               unsigned nblocks, int identity_vals) {
    __shared__ unsigned wave_hist[kRadixWaves][kRadix];
    __shared__ unsigned base[kRadixWaves][kRadix];
    const unsigned t = threadIdx.x, lane = t & 63u, w = t >> 6;
    for (unsigned k = t; k < kRadixWaves * kRadix; k += kBuildBlock)
        (&wave_hist[0][0])[k] = 0;
    __syncthreads();
    const unsigned sub = blockIdx.x * kRadixTile + w * (kRadixRounds * 64);
    for (unsigned r = 0; r < kRadixRounds; ++r) {
        const unsigned idx = sub + r * 64 + lane;
        if (idx < n)
            atomicAdd(&wave_hist[w][(keys_in[idx] >> shift) & (kRadix - 1)], 1u);
    }
    __syncthreads();
    {
        unsigned run = offsets[t * nblocks + blockIdx.x];
        for (unsigned ww = 0; ww < kRadixWaves; ++ww) {
            base[ww][t] = run;
            run += wave_hist[ww][t];
        }
    }
    __syncthreads();
    const unsigned long long below = (1ull << lane) - 1ull;
    for (unsigned r = 0; r < kRadixRounds; ++r) {
        const unsigned idx = sub + r * 64 + lane;
        const bool valid = idx < n;
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
            const unsigned pos = base[w][d] + rank;
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __builtin_amdgcn_wave_barrier();
        if (valid && rank == 0)
            base[w][d] += __popcll(m);
        __builtin_amdgcn_wave_barrier();
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
    const unsigned key = keys[i];
    const bool first = i == 0 || keys[i - 1] != key;
    const bool last = i + 1 == n || keys[i + 1] != key;
    if (first || last) {
        TayCellRange *r = ranges + key_to_cell(key, dims, g);
        if (first)
            r->beg = i;
        if (last)
            r->end = i + 1;
    }
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
    dst[i] = key;
    dst[cap + i] = i - ranges[key_to_cell(key, dims, g)].beg;
    for (unsigned w = 2; w < words; ++w)
        dst[w * cap + i] = src[w * cap + from];
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
    return hip_ok(hipGetLastError(), what);
}

}  // namespace

/* -------------------------------------------------------------- C-ABI */

TAY_API int tay_gpu_available(void) {
    return stream() != nullptr;
}

TAY_API const char *tay_gpu_last_error(void) {
    return g_last_error.c_str();
}

TAY_API int tay_gpu_synchronize(void) {
    hipStream_t s = stream();
    return s && hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");
}

TAY_API void tay_gpu_space_destroy(TayGpuSpace *sp) {
    if (!sp)
        return;
    if (g_device_state == 1)
        (void)hipStreamSynchronize(g_stream);
    device_free(sp->cols[0]);
    device_free(sp->cols[1]);
    device_free(sp->geom);
    device_free(sp->box_partials);
    device_free(sp->keys[0]);
    device_free(sp->keys[1]);
    device_free(sp->vals[0]);
    device_free(sp->vals[1]);
    device_free(sp->block_hist);
    device_free(sp->scan_sums);
    device_free(sp->ranges);
    device_free(sp->staging);
    delete sp;
}

TAY_API TayGpuSpace *tay_gpu_space_create(int kind, unsigned agent_size, unsigned capacity, int is_point, int dims,
                                          const float *part_sizes, unsigned long long shared_bytes) {
    if (kind != TAY_GPU_KIND_Z_GRID && kind != TAY_GPU_KIND_SIMPLE) {
        set_error("unknown GPU space kind %d", kind);
        return nullptr;
    }
    /* the tag (8 bytes) and the position (non-point: a float4 min corner,
     * then the max corner) over the used axes */
    const unsigned min_size = (is_point ? 8u : 24u) + 4u * (unsigned)(dims > 0 ? dims : 1);
    if (agent_size < min_size || agent_size % 4 != 0) {
        set_error("agent size %u must be a multiple of 4 of at least %u bytes (tag and position)", agent_size,
                  min_size);
        return nullptr;
    }
    if (dims < 1 || dims > 4 || (kind == TAY_GPU_KIND_Z_GRID && dims > 3)) {
        set_error("%d dimensions not supported by this space (Z-grid: 1 to 3)", dims);
        return nullptr;
    }
    if (kind == TAY_GPU_KIND_Z_GRID && !is_point) {
        set_error("grid structure can only contain point agents");
        return nullptr;
    }
    if (!stream())
        return nullptr;

    TayGpuSpace *sp = new TayGpuSpace();
    std::memset(sp, 0, sizeof(*sp));
    sp->kind = kind;
    sp->agent_size = agent_size;
    sp->words = agent_size / 4;
    sp->capacity = capacity ? capacity : 1;
    sp->is_point = is_point;
    sp->dims = dims;
    for (int k = 0; k < 4; ++k)
        sp->part[k] = part_sizes ? part_sizes[k] : 1.0f;
    sp->shared_bytes = shared_bytes;

    const size_t col_bytes = (size_t)sp->words * sp->capacity * 4;
    bool ok = device_alloc((void **)&sp->cols[0], col_bytes, "agent columns") &&
              device_alloc((void **)&sp->cols[1], col_bytes, "agent columns");
    if (ok && kind == TAY_GPU_KIND_Z_GRID) {
        for (int k = 0; k < dims; ++k)
            if (!(sp->part[k] > 0.0f)) {
                set_error("partition size of axis %d must be positive", k);
                ok = false;
            }
        sp->max_cells = shared_bytes / sizeof(TayCellRange);
        if (ok && sp->max_cells == 0) {
            set_error("Z-grid needs shared memory for its cell table (shared_size_in_megabytes > 0)");
            ok = false;
        }
        sp->box_blocks = blocks_for(sp->capacity, kBuildBlock * 4, 1024);
        sp->radix_blocks = (sp->capacity + kRadixTile - 1) / kRadixTile;
        ok = ok && device_alloc((void **)&sp->geom, sizeof(TayGridGeom), "grid geometry") &&
             device_alloc((void **)&sp->box_partials, sizeof(float) * 6 * sp->box_blocks, "box partials") &&
             device_alloc((void **)&sp->keys[0], 4ull * sp->capacity, "sort keys") &&
             device_alloc((void **)&sp->keys[1], 4ull * sp->capacity, "sort keys") &&
             device_alloc((void **)&sp->vals[0], 4ull * sp->capacity, "sort values") &&
             device_alloc((void **)&sp->vals[1], 4ull * sp->capacity, "sort values") &&
             device_alloc((void **)&sp->block_hist, 4ull * kRadix * (sp->radix_blocks ? sp->radix_blocks : 1),
                          "radix histogram") &&
             device_alloc((void **)&sp->scan_sums, 4ull * kScanBlocks, "scan sums") &&
             device_alloc((void **)&sp->ranges, sizeof(TayCellRange) * sp->max_cells, "cell table");
        if (ok)
            ok = hip_ok(hipMemsetAsync(sp->geom, 0, sizeof(TayGridGeom), stream()), "hipMemsetAsync");
    }
    if (!ok) {
        tay_gpu_space_destroy(sp);
        return nullptr;
    }
    return sp;
}

TAY_API int tay_gpu_space_fix_box(TayGpuSpace *sp, const float *min, const float *max) {
    for (int k = 0; k < 4; ++k) {
        sp->fixed_min[k] = min[k];
        sp->fixed_max[k] = max[k];
    }
    sp->box_fixed = 1;
    return 1;
}

TAY_API int tay_gpu_space_unfix_box(TayGpuSpace *sp) {
    sp->box_fixed = 0;
    return 1;
}

TAY_API unsigned tay_gpu_space_count(const TayGpuSpace *sp) {
    return sp->count;
}

static bool ensure_staging(TayGpuSpace *sp, size_t bytes) {
    if (sp->staging_bytes >= bytes)
        return true;
    device_free(sp->staging);
    sp->staging = nullptr;
    sp->staging_bytes = 0;
    if (!device_alloc((void **)&sp->staging, bytes, "transfer staging"))
        return false;
    sp->staging_bytes = bytes;
    return true;
}

TAY_API int tay_gpu_space_push(TayGpuSpace *sp, const void *agents, unsigned count) {
    if (count > sp->capacity) {
        set_error("%u agents exceed the group capacity %u", count, sp->capacity);
        return 0;
    }
    hipStream_t s = stream();
    if (!s)
        return 0;
    sp->count = count;
    if (count == 0)
        return 1;
    const size_t bytes = (size_t)count * sp->agent_size;
    if (!ensure_staging(sp, bytes) ||
        !hip_ok(hipMemcpyAsync(sp->staging, agents, bytes, hipMemcpyHostToDevice, s), "push agents"))
        return 0;
    k_aos_to_soa<<<blocks_for((unsigned long long)count * sp->words, kBuildBlock, 8192), kBuildBlock, 0, s>>>(
        sp->staging, sp->cols[sp->cur], count, sp->words, sp->capacity);
    return launch_ok("aos_to_soa");
}

TAY_API int tay_gpu_space_fetch(TayGpuSpace *sp, void *agents, unsigned count) {
    hipStream_t s = stream();
    if (!s)
        return 0;
    if (count > sp->count)
        count = sp->count;
    if (count == 0)
        return 1;
    const size_t bytes = (size_t)count * sp->agent_size;
    if (!ensure_staging(sp, bytes))
        return 0;
    k_soa_to_aos<<<blocks_for((unsigned long long)count * sp->words, kBuildBlock, 8192), kBuildBlock, 0, s>>>(
        sp->cols[sp->cur], sp->staging, count, sp->words, sp->capacity);
    if (!launch_ok("soa_to_aos"))
        return 0;
    return hip_ok(hipMemcpyAsync(agents, sp->staging, bytes, hipMemcpyDeviceToHost, s), "fetch agents") &&
           hip_ok(hipStreamSynchronize(s), "fetch agents");
}

TAY_API int tay_gpu_space_sort(TayGpuSpace *sp) {
    if (sp->kind != TAY_GPU_KIND_Z_GRID || sp->count == 0)
        return 1;
    hipStream_t s = stream();
    if (!s)
        return 0;
    const unsigned n = sp->count, cap = sp->capacity;
    const unsigned *src = sp->cols[sp->cur];
    unsigned *dst = sp->cols[sp->cur ^ 1];

    const unsigned box_blocks = blocks_for(n, kBuildBlock * 4, sp->box_blocks);
    k_box_partial<<<box_blocks, kBuildBlock, 0, s>>>(src, cap, n, sp->dims, sp->is_point, sp->box_partials);
    k_geometry<<<1, 64, 0, s>>>(sp->box_partials, box_blocks, sp->dims, sp->part[0], sp->part[1], sp->part[2],
                                sp->max_cells, sp->box_fixed, sp->fixed_min[0], sp->fixed_min[1],
                                sp->fixed_min[2], sp->fixed_max[0], sp->fixed_max[1], sp->fixed_max[2], sp->geom);
    const unsigned agent_blocks = blocks_for(n, kBuildBlock, 0xFFFFFFFFu);
    k_keys<<<agent_blocks, kBuildBlock, 0, s>>>(src, cap, n, sp->dims, sp->geom, sp->keys[0]);

    const unsigned nblocks = (n + kRadixTile - 1) / kRadixTile;
    for (unsigned pass = 0; pass < kRadixPasses; ++pass) {
        const unsigned shift = pass * kRadixBits;
        const unsigned in = pass & 1u, out = in ^ 1u;
        const unsigned total = kRadix * nblocks;
        const unsigned segs = total < kScanBlocks * kBuildBlock ? (total + kBuildBlock - 1) / kBuildBlock : kScanBlocks;
        k_radix_hist<<<nblocks, kBuildBlock, 0, s>>>(sp->keys[in], n, shift, sp->block_hist, nblocks);
        k_scan_reduce<<<segs, kBuildBlock, 0, s>>>(sp->block_hist, total, sp->scan_sums);
        k_scan_top<<<1, 64, 0, s>>>(sp->scan_sums, segs);
        k_scan_down<<<segs, kBuildBlock, 0, s>>>(sp->block_hist, total, sp->scan_sums);
        k_radix_scatter<<<nblocks, kBuildBlock, 0, s>>>(sp->keys[in], sp->vals[in], sp->keys[out], sp->vals[out], n,
                                                        shift, sp->block_hist, nblocks, pass == 0);
    }
    /* kRadixPasses is even: the sorted keys and source indices are back in buffer 0 */
    k_ranges_clear<<<blocks_for(sp->max_cells, kBuildBlock, 4096), kBuildBlock, 0, s>>>(sp->ranges, sp->geom);
    k_ranges<<<agent_blocks, kBuildBlock, 0, s>>>(sp->keys[0], n, sp->dims, sp->geom, sp->ranges);
    k_permute<<<agent_blocks, kBuildBlock, 0, s>>>(sp->keys[0], sp->vals[0], n, sp->words, cap, sp->dims, sp->geom,
                                                   sp->ranges, src, dst);
    if (!launch_ok("z-grid sort"))
        return 0;
    sp->cur ^= 1;
    return 1;
}

TAY_API int tay_gpu_copy_probe(unsigned long long bytes, double *ms) {
    hipStream_t s = stream();
    if (!s)
        return 0;
    bytes &= ~15ull;
    void *src = nullptr, *dst = nullptr;
    hipEvent_t beg = nullptr, end = nullptr;
    float t = 0.0f;
    bool ok = device_alloc(&src, bytes, "probe source") && device_alloc(&dst, bytes, "probe destination") &&
              hip_ok(hipMemsetAsync(src, 0, bytes, s), "hipMemsetAsync") &&
              hip_ok(hipEventCreate(&beg), "hipEventCreate") && hip_ok(hipEventCreate(&end), "hipEventCreate") &&
              hip_ok(hipEventRecord(beg, s), "hipEventRecord");
    if (ok) {
        k_copy_probe<<<2048, kBuildBlock, 0, s>>>((const uint4 *)src, (uint4 *)dst, bytes / 16);
        ok = launch_ok("copy probe") && hip_ok(hipEventRecord(end, s), "hipEventRecord") &&
             hip_ok(hipEventSynchronize(end), "hipEventSynchronize") &&
             hip_ok(hipEventElapsedTime(&t, beg, end), "hipEventElapsedTime");
    }
    if (beg)
        (void)hipEventDestroy(beg);
    if (end)
        (void)hipEventDestroy(end);
    device_free(src);
    device_free(dst);
    if (ms)
        *ms = t;
    return ok;
}

TAY_API int tay_gpu_space_grid_info(TayGpuSpace *sp, float *origin, float *cell_sizes, int *cell_counts,
                                    unsigned long long *cells_count) {
    if (sp->kind != TAY_GPU_KIND_Z_GRID) {
        set_error("not a Z-grid space");
        return 0;
    }
    TayGridGeom g;
    hipStream_t s = stream();
    if (!s || !hip_ok(hipMemcpyAsync(&g, sp->geom, sizeof(g), hipMemcpyDeviceToHost, s), "grid info") ||
        !hip_ok(hipStreamSynchronize(s), "grid info"))
        return 0;
    for (int k = 0; k < 4; ++k) {
        origin[k] = g.origin[k];
        cell_sizes[k] = g.cell_size[k];
        cell_counts[k] = g.count[k];
    }
    *cells_count = g.cells;
    return 1;
}
