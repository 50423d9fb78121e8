/* internal.h - declarations shared by the translation units of libtay_amd. */
#ifndef TAY_INTERNAL_H
#define TAY_INTERNAL_H

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "device_abi.h"
#include "tay_gpu.h"

#define TAY_API extern "C" __attribute__((visibility("default")))

namespace tay {

/* Records a printf-style message as the calling thread's last error. */
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

/* Returns false (with the error recorded) when a HIP call failed. */
bool hip_ok(hipError_t e, const char *what);

/* The library's stream, created on first use on device 0.  Returns nullptr
 * (error recorded) when no GPU can be used. */
hipStream_t stream();

/* Device allocation helpers; false with the error recorded on failure. */
bool device_alloc(void **ptr, size_t bytes, const char *what);
/* Sum of a see pass's TAY_PAIR_SLOTS device pair counters (jit.cpp). */
int read_pair_counters(const unsigned long long *counters, unsigned long long *pairs);
void device_free(void *ptr);

/* Sub-cell bits per axis of the storage order inside a cell (0: previous
 * storage order only). */
constexpr int TAY_GPU_DEFAULT_ORDER_BITS = 2;

/* Widest lane group of a see tile that walks per lane when its box is too
 * wide for one shared sweep (0: split down to single lanes instead). */
constexpr unsigned TAY_GPU_DEFAULT_LANE_WALK = 64;

constexpr unsigned kSeeBlock = 256;     /* 4 wavefronts, one 64-seer tile each */
constexpr unsigned kActBlock = 256;

/* hipRTC compile of a composed program for gfx950 (cached by text); on a
 * compile error the log goes to log. */
bool compile_program(const std::string &text, std::vector<char> &code, char *log, unsigned log_size);

/* Program text of a batched ensemble: model sources and one
 * tay_batch_pass_<i> kernel per pass. */
std::string compose_batch(const char *const *sources, unsigned sources_count, const TayGpuBatchPass *passes,
                          unsigned passes_count, unsigned agent_size, int dims, int count_pairs);

/* True for a C identifier. */
bool valid_func_name(const char *name);

/* Text of device_abi.h and jit_device.h, embedded by the build. */
extern const char *const device_abi_text;
extern const char *const device_math_text;
extern const char *const jit_device_text;

}  // namespace tay

struct TayGpuSpace {
    int kind;
    unsigned agent_size;
    unsigned words;             /* 32-bit words per agent */
    unsigned capacity;
    int is_point;
    int dims;
    float part[4];
    int order_bits;             /* sub-cell bits of the storage order */
    unsigned long long shared_bytes;
    unsigned long long max_cells;
    unsigned count;

    unsigned *cols[2];          /* structure-of-arrays agent storage, ping-pong */
    int cur;

    /* Z-grid build */
    TayGridGeom *geom;
    float *box_partials;
    unsigned box_blocks;
    unsigned *keys[2];
    unsigned *vals[2];
    unsigned *block_hist;
    unsigned *scan_sums;
    unsigned radix_blocks;
    TayCellRange *ranges;
    int box_fixed;
    float fixed_min[4];
    float fixed_max[4];

    /* host <-> device transfers */
    unsigned *staging;
    size_t staging_bytes;
};

struct TayGpuProgram {
    struct Pass {
        TayGpuPassDesc desc;
        std::string func_name;
        std::string split_clear, split_combine;     /* copies behind desc.split_clear / split_combine */
        hipFunction_t kernel;
        void *context;              /* device copy of the pass context */
        void *context_host;         /* pinned staging for the upload */
        unsigned long long *pair_count;
        /* timing: event pairs recorded around launches, resolved on query */
        std::vector<hipEvent_t> ev_free;
        std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pending;
        double ms;
        unsigned launches;
    };
    hipModule_t module;
    std::vector<Pass> passes;
    int count_pairs;
    int timing;
    unsigned width;             /* seers per lane group at the start of a see tile */
    unsigned lane_walk;         /* sparse lane groups up to this width walk per lane */
    unsigned tile_seers;        /* seers per see tile: 64, 32 or 16; 0 chooses by the seer count */
};

#endif
