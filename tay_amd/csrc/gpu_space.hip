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
#include "device_math.h"

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

#include "grid_build.h"

/* -------------------------------------------------------------- C-ABI */

TAY_API int tay_gpu_available(void) {
    return stream() != nullptr;
}

/* Page-locked host memory is handed out up to this size per allocation;
 * larger host storage stays pageable. */
constexpr size_t kMaxPinnedBytes = 4ull << 30;

TAY_API void *tay_gpu_host_alloc(size_t bytes) {
    if (bytes == 0 || bytes > kMaxPinnedBytes || !stream())
        return nullptr;
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    std::memset(p, 0, bytes);
    return p;
}

TAY_API void tay_gpu_host_free(void *p) {
    if (p)
        (void)hipHostFree(p);
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
    /* kernels index a column word as word * capacity + slot, and a slot's
     * byte offset in a column, in 32 bits */
    if ((unsigned long long)(agent_size / 4) * (capacity ? capacity : 1) > 0xFFFFFFFFull) {
        set_error("capacity %u of %u-byte agents exceeds 2^32 column words", capacity, agent_size);
        return nullptr;
    }
    if (capacity > (1u << 30)) {
        set_error("capacity %u exceeds 2^30 agents (32-bit byte offsets in a column)", capacity);
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
    sp->order_bits = TAY_GPU_DEFAULT_ORDER_BITS;

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
             device_alloc((void **)&sp->box_partials, sizeof(float) * kBoxPartial * sp->box_blocks, "box partials") &&
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

TAY_API int tay_gpu_space_set_order_bits(TayGpuSpace *sp, int bits) {
    if (bits < 0 || bits > 3) {
        set_error("order bits %d: must be 0 to 3", bits);
        return 0;
    }
    sp->order_bits = bits;
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
    /* from page-locked storage the copy runs asynchronously: wait for it,
     * since the caller may change the records as soon as this returns */
    return launch_ok("aos_to_soa") && hip_ok(hipStreamSynchronize(s), "push agents");
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

TAY_API int tay_gpu_space_fetch_words(TayGpuSpace *sp, unsigned first_word, unsigned words, void *dst,
                                     unsigned count) {
    if (first_word > sp->words || words > sp->words - first_word) {
        set_error("words [%u, %u + %u) outside the %u words of an agent", first_word, first_word, words, sp->words);
        return 0;
    }
    hipStream_t s = stream();
    if (!s)
        return 0;
    if (count > sp->count)
        count = sp->count;
    if (count == 0 || words == 0)
        return 1;
    const size_t bytes = (size_t)count * words * 4u;
    if (!ensure_staging(sp, bytes))
        return 0;
    k_soa_to_aos<<<blocks_for((unsigned long long)count * words, kBuildBlock, 8192), kBuildBlock, 0, s>>>(
        sp->cols[sp->cur] + (size_t)first_word * sp->capacity, sp->staging, count, words, sp->capacity);
    if (!launch_ok("soa_to_aos"))
        return 0;
    return hip_ok(hipMemcpyAsync(dst, sp->staging, bytes, hipMemcpyDeviceToHost, s), "fetch words") &&
           hip_ok(hipStreamSynchronize(s), "fetch words");
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

    /* A small space (one block sorts it) has its sort write the cell ranges
     * too: the chain of a small step is bound by kernel-to-kernel latency,
     * not by its bytes.  (Its box stays a many-block reduction: one block
     * reading 10,000 positions took 20 us against 5 + 7 us for the box and
     * geometry kernels.) */
    const bool small = n <= kSmallSort;
    const unsigned box_blocks = blocks_for(n, kBuildBlock * 4, sp->box_blocks);
    k_box_partial<<<box_blocks, kBuildBlock, 0, s>>>(src, cap, n, sp->dims, sp->is_point, sp->box_partials);
    k_geometry<<<1, 64, 0, s>>>(sp->box_partials, box_blocks, sp->dims, sp->part[0], sp->part[1], sp->part[2],
                                sp->max_cells, sp->order_bits, sp->box_fixed, sp->fixed_min[0], sp->fixed_min[1],
                                sp->fixed_min[2], sp->fixed_max[0], sp->fixed_max[1], sp->fixed_max[2], sp->geom);
    const unsigned agent_blocks = blocks_for(n, kBuildBlock, 0xFFFFFFFFu);
    const unsigned clear_blocks = blocks_for(sp->max_cells, kBuildBlock, 4096);
    k_keys_clear<<<agent_blocks > clear_blocks ? agent_blocks : clear_blocks, kBuildBlock, 0, s>>>(
        src, cap, n, sp->dims, sp->geom, sp->keys[0], sp->ranges);

    const unsigned nblocks = (n + kRadixTile - 1) / kRadixTile;
    if (small) {
        k_small_sort<<<1, kRadixBlock, 0, s>>>(sp->keys[0], sp->vals[0], n, sp->dims, sp->geom, sp->ranges);
    } else {
        for (unsigned pass = 0; pass < kRadixPasses; ++pass) {
            const unsigned shift = pass * kRadixBits;
            const unsigned in = pass & 1u, out = in ^ 1u;
            const unsigned total = kRadix * nblocks;
            const unsigned segs =
                total < kScanBlocks * kBuildBlock ? (total + kBuildBlock - 1) / kBuildBlock : kScanBlocks;
            k_radix_hist<<<nblocks, kRadixBlock, 0, s>>>(sp->keys[in], n, shift, sp->block_hist, nblocks, nullptr,
                                                         nullptr);
            k_scan_reduce<<<segs, kBuildBlock, 0, s>>>(sp->block_hist, total, sp->scan_sums, nullptr, shift);
            k_scan_down<<<segs, kBuildBlock, 0, s>>>(sp->block_hist, total, sp->scan_sums, nullptr, shift);
            k_radix_scatter<<<nblocks, kRadixBlock, 0, s>>>(sp->keys[in], sp->vals[in], sp->keys[out], sp->vals[out],
                                                            n, shift, sp->block_hist, nblocks, pass == 0, nullptr,
                                                            nullptr);
        }
    }
    /* kRadixPasses is even: the sorted keys and source indices are back in
     * buffer 0 (the small sort works in place there and writes the ranges) */
    if (!small)
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

/* Every float bit pattern x in [first, first + count): tay_sqrt_rn(x)
 * against the compiler's IEEE float square root and against the double
 * square root rounded to float (exact for square roots: 53 >= 2*24 + 2
 * bits), counting bitwise mismatches; the first mismatching input of each
 * comparison is kept (atomicMin). */
__global__ __launch_bounds__(kBuildBlock) void k_sqrt_selftest(unsigned long long first, unsigned long long count,
                                                               unsigned long long *counts, unsigned *first_bad) {
    unsigned long long bad_f = 0, bad_d = 0;
    unsigned min_f = 0xFFFFFFFFu, min_d = 0xFFFFFFFFu;
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long k = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; k < count; k += stride) {
        const unsigned bits = (unsigned)(first + k);
        const float x = __uint_as_float(bits);
        const unsigned got = __float_as_uint(tay_sqrt_rn(x));
        const unsigned want_f = __float_as_uint(__builtin_sqrtf(x));
        const unsigned want_d = __float_as_uint((float)__builtin_sqrt((double)x));
        if (got != want_f) {
            ++bad_f;
            min_f = bits < min_f ? bits : min_f;
        }
        if (got != want_d) {
            ++bad_d;
            min_d = bits < min_d ? bits : min_d;
        }
    }
    if (bad_f) {
        atomicAdd(&counts[0], bad_f);
        atomicMin(&first_bad[0], min_f);
    }
    if (bad_d) {
        atomicAdd(&counts[1], bad_d);
        atomicMin(&first_bad[1], min_d);
    }
}

TAY_API int tay_gpu_sqrt_selftest(unsigned long long first, unsigned long long count,
                                  unsigned long long *mismatches, unsigned *first_bad) {
    hipStream_t s = stream();
    if (!s)
        return 0;
    if (first > 0xFFFFFFFFull || count > 0x100000000ull - first) {
        set_error("sqrt self-test range [%llu, %llu + %llu) leaves the 32-bit patterns", first, first, count);
        return 0;
    }
    void *buf = nullptr;
    const unsigned init[4] = {0u, 0u, 0xFFFFFFFFu, 0xFFFFFFFFu};   /* counts[2] as 4 words, first_bad[2] */
    bool ok = device_alloc(&buf, 32, "sqrt self-test counters") &&
              hip_ok(hipMemsetAsync(buf, 0, 16, s), "hipMemsetAsync") &&
              hip_ok(hipMemcpyAsync((char *)buf + 16, init + 2, 8, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    if (ok && count) {
        k_sqrt_selftest<<<4096, kBuildBlock, 0, s>>>(first, count, (unsigned long long *)buf,
                                                     (unsigned *)((char *)buf + 16));
        ok = launch_ok("sqrt self-test");
    }
    unsigned long long host[3] = {0, 0, 0};
    ok = ok && hip_ok(hipMemcpyAsync(host, buf, 24, hipMemcpyDeviceToHost, s), "hipMemcpyAsync") &&
         hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");
    device_free(buf);
    if (!ok)
        return 0;
    if (mismatches) {
        mismatches[0] = host[0];
        mismatches[1] = host[1];
    }
    if (first_bad) {
        first_bad[0] = (unsigned)(host[2] & 0xFFFFFFFFu);
        first_bad[1] = (unsigned)(host[2] >> 32);
    }
    return 1;
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
