/*
 * gpu_static.hip - the allocation-free variant of tay_gpu_space_fetch_words.
 *
 * tay_gpu_space_fetch_words stages the selected words in a device buffer of
 * the space, allocated (hipMalloc) the first time a fetch needs it and again
 * whenever a larger one does.  tay_gpu_space_fetch_words_static stages them
 * in `scratch`, a file-scope device array of kStaticScratchBytes that every
 * GPU holds from the library's load until the process ends, and calls no
 * allocating routine: for callers that cannot allocate (code under HIP graph
 * capture, tight viewer loops, harnesses that forbid it).
 *
 * In host code the array's name is a host-side stand-in, not a device
 * address: its device address is looked up once per device with
 * hipGetSymbolAddress (with that device current) and kept; every use below
 * goes through that pointer.
 */
#include "internal.h"

#include <cstring>

using namespace tay;

#include "grid_build.h"

namespace {

/* 32 MiB: the positions (12 bytes) of 2.7M agents, the positions and
 * velocities of 1.3M. */
constexpr size_t kStaticScratchBytes = 32u << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

void *g_scratch_ptr[kMaxDevices];   /* device address of `scratch` per device, once looked up */

/* The calling thread's current device's address of `scratch`, looked up on
 * the first call on that device. */
void *scratch_ptr() {
    int dev = 0;
    if (!hip_ok(hipGetDevice(&dev), "hipGetDevice"))
        return nullptr;
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device %d beyond the %d devices the static scratch table covers", dev, kMaxDevices);
        return nullptr;
    }
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        if (!hip_ok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)"))
            return nullptr;
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

}  // namespace

TAY_API int tay_gpu_space_fetch_words_static(TayGpuSpace *sp, unsigned first_word, unsigned words, void *dst,
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
    const unsigned long long bytes = (unsigned long long)count * words * 4u;
    if (bytes > kStaticScratchBytes) {
        set_error("%u agents x %u words = %llu bytes exceed the %zu-byte static staging buffer", count, words,
                  bytes, kStaticScratchBytes);
        return 0;
    }
    unsigned *staging = (unsigned *)scratch_ptr();
    if (!staging)
        return 0;
    /* every staged word is written before it is copied out: nothing needs
     * clearing first */
    k_soa_to_aos<<<blocks_for((unsigned long long)count * words, kBuildBlock, 8192), kBuildBlock, 0, s>>>(
        sp->cols[sp->cur] + (size_t)first_word * sp->capacity, staging, count, words, sp->capacity);
    if (!launch_ok("soa_to_aos (static staging)"))
        return 0;
    return hip_ok(hipMemcpyAsync(dst, staging, (size_t)bytes, hipMemcpyDeviceToHost, s), "fetch words (static)") &&
           hip_ok(hipStreamSynchronize(s), "fetch words (static)");
}
