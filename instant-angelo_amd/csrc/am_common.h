// Shared device/host helpers for the MI355X (gfx950) NeuS hot-path library.
#pragma once
#include <hip/hip_runtime.h>
#include <mutex>
#include <stdint.h>
#include <stddef.h>
#include "../../include/angelo_amd.h"

namespace am {

// ---- error reporting (thread-local message, int status through the C-ABI) ----
void set_error(const char* fmt, ...);
const char* last_error();
// Records the workspace bytes a call needs (am_last_workspace_need) and the
// message; returns AM_ERR_WORKSPACE.  `used`: the carver's bytes.
int workspace_short(size_t used, size_t got);
size_t last_workspace_need();

#define AM_CHECK_ARG(cond, ...)                 \
    do {                                        \
        if (!(cond)) {                          \
            ::am::set_error(__VA_ARGS__);       \
            return AM_ERR_INVALID;              \
        }                                       \
    } while (0)

#define AM_HIP(call)                                                              \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                   \
            ::am::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                            __FILE__, __LINE__);                                  \
            return AM_ERR_HIP;                                                    \
        }                                                                         \
    } while (0)

#define AM_LAUNCH_CHECK()                                                               \
    do {                                                                                \
        hipError_t e_ = hipGetLastError();                                              \
        if (e_ != hipSuccess) {                                                         \
            ::am::set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(e_), \
                            __FILE__, __LINE__);                                        \
            return AM_ERR_HIP;                                                          \
        }                                                                               \
    } while (0)

static inline hipStream_t to_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Kernel timers (am_timer_enable / am_timer_read): a scope records a HIP
// event pair around the launches it encloses when timing is on.
bool timer_on();
void timer_push(const char* name, hipEvent_t a, hipEvent_t b);
// Adds work units (e.g. table-gradient pairs) under `name` while timers are on.
void timer_units(const char* name, int64_t units);
// Implementation variant switches set by am_set_flag (name -> value), read
// on the host at launch time; `dflt` when the flag was never set.
int flag_value(const char* name, int dflt);
hipEvent_t timer_event();
// While set, timed launches are booked under this name instead of their own
// (an entry point reusing another's kernels, e.g. the occupancy update).
const char* timer_alias();
void set_timer_alias(const char* alias);

struct TimerAlias {
    const char* prev;
    explicit TimerAlias(const char* a) : prev(timer_alias()) { set_timer_alias(a); }
    ~TimerAlias() { set_timer_alias(prev); }
};

struct KernelTimer {
    const char* name;
    hipStream_t s;
    hipEvent_t a = nullptr, b = nullptr;
    KernelTimer(const char* n, hipStream_t st) : name(timer_alias() ? timer_alias() : n), s(st) {
        if (timer_on()) {
            a = timer_event();
            b = timer_event();
            if (a && b) hipEventRecord(a, s);
        }
    }
    ~KernelTimer() {
        if (a && b) {
            hipEventRecord(b, s);
            timer_push(name, a, b);
        }
    }
};
static inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
static inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Carves aligned sub-buffers out of one caller-provided workspace.
struct WsCarver {
    char* base;
    size_t cap, used;
    WsCarver(void* p, size_t c) : base(static_cast<char*>(p)), cap(c), used(0) {}
    template <typename T>
    T* take(size_t count) {
        used = align_up(used, 256);
        T* r = base ? reinterpret_cast<T*>(base + used) : nullptr;
        used += count * sizeof(T);
        return r;
    }
    bool ok() const { return used <= cap; }
};

typedef float v4f __attribute__((ext_vector_type(4)));

// Deterministic column sum out[c] = sum_r in[r*cols + c] (r in increasing
// order within fixed slices, slices summed in order).  tmp: workspace of
// colsum_tmp_floats(rows, cols) floats.
int64_t colsum_tmp_floats(int64_t rows, int64_t cols);
int colsum_f32(const float* in, int64_t rows, int64_t cols, float* out, float* tmp,
               hipStream_t stream);

// Workgroups are dispatched to the chip's 8 XCDs round-robin (block b runs
// on XCD b % 8, each XCD with its own L2).  Renumbering b to
// start(b % 8) + b / 8 gives each XCD a consecutive range of virtual block
// indices, so work items that neighbour each other in index order (samples
// along one ray) share an L2.  A bijection of [0, g); identity below 8 blocks.
__device__ __forceinline__ int xcd_block(int b, int g) {
    if (g < 8) return b;
    const int x = b & 7, j = b >> 3;
    const int q = g >> 3, r = g & 7;
    return x * q + (x < r ? x : r) + j;
}

// Persistent-grid block count: enough blocks for `tiles` work items of
// `per_block` each, capped at `max_blocks`.
static inline int persistent_blocks(int64_t tiles, int per_block, int max_blocks) {
    int64_t b = (tiles + per_block - 1) / per_block;
    if (b > max_blocks) b = max_blocks;
    if (b < 1) b = 1;
    return (int)b;
}

// Blocks of `block` threads with `lds` bytes of dynamic LDS that the device
// holds at once (occupancy per CU x CUs), for kernels whose register count
// leaves wave slots a fixed-size grid would not fill in one round.  Cached
// per (kernel, LDS size); at least 1.
static inline int resident_blocks(const void* kern, int block, size_t lds) {
    struct Entry { const void* k; size_t lds; int blocks; };
    static Entry cache[16];
    static int n_cached = 0;
    static std::mutex mu;
    std::lock_guard<std::mutex> lock(mu);
    for (int i = 0; i < n_cached; ++i)
        if (cache[i].k == kern && cache[i].lds == lds) return cache[i].blocks;
    int per_cu = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, block, lds) != hipSuccess) per_cu = 1;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 1;
    const int blocks = (per_cu < 1 ? 1 : per_cu) * (cus < 1 ? 1 : cus);
    if (n_cached < 16) cache[n_cached++] = Entry{kern, lds, blocks};
    return blocks;
}

// ---- wave helpers (wave64) ----
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

__device__ __forceinline__ float shfl(float v, int src_lane) {
    return __shfl(v, src_lane, 64);
}

// Sum over the four lane groups (lanes l, l^16, l^32, l^48).
__device__ __forceinline__ float sum_groups(float v) {
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
}

// Sum over the 16 lanes of a lane group (lanes sharing l/16).
__device__ __forceinline__ float sum_in_group(float v) {
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    v += __shfl_xor(v, 8, 64);
    return v;
}

__device__ __forceinline__ v4f mfma4(float a, float b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

}  // namespace am
