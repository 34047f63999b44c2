// Error reporting, version, kernel timers and the deterministic column sum.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "am_common.h"
#include <mutex>
#include <string>
#include <utility>
#include <vector>

namespace am {

struct TimedRecord {
    std::string name;
    hipEvent_t a, b;
};
static bool g_timer_on = false;
static std::vector<TimedRecord> g_records;
struct UnitRecord {
    std::string name;
    int64_t units;
};
static std::vector<UnitRecord> g_units;
static std::vector<hipEvent_t> g_event_pool;

bool timer_on() { return g_timer_on; }

static const char* g_timer_alias = nullptr;
const char* timer_alias() { return g_timer_alias; }
void set_timer_alias(const char* a) { g_timer_alias = a; }

hipEvent_t timer_event() {
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void timer_push(const char* name, hipEvent_t a, hipEvent_t b) {
    g_records.push_back(TimedRecord{name, a, b});
}

void timer_units(const char* name, int64_t units) {
    if (!g_timer_on) return;
    for (auto& u : g_units)
        if (u.name == name) {
            u.units += units;
            return;
        }
    g_units.push_back(UnitRecord{name, units});
}

static void timer_reset() {
    g_units.clear();
    for (auto& r : g_records) {
        g_event_pool.push_back(r.a);
        g_event_pool.push_back(r.b);
    }
    g_records.clear();
}

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

const char* last_error() { return g_err; }

static thread_local size_t g_ws_need = 0;

int workspace_short(size_t used, size_t got) {
    g_ws_need = used + 256;   // the size queries' alignment slack
    set_error("workspace too small: need %zu bytes, got %zu", g_ws_need, got);
    return AM_ERR_WORKSPACE;
}

size_t last_workspace_need() { return g_ws_need; }

static constexpr int64_t kColsumSlice = 64;  // rows per slice

int64_t colsum_tmp_floats(int64_t rows, int64_t cols) {
    const int64_t slices = (rows + kColsumSlice - 1) / kColsumSlice;
    return slices > 1 ? slices * cols : 0;
}

__global__ void k_colsum(const float* __restrict__ in, int64_t rows, int64_t cols, int64_t slice,
                         float* __restrict__ out) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    const int64_t r0 = (int64_t)blockIdx.y * slice;
    const int64_t r1 = r0 + slice < rows ? r0 + slice : rows;
    float s = 0.0f;
    for (int64_t r = r0; r < r1; ++r) s += in[r * cols + c];
    out[(int64_t)blockIdx.y * cols + c] = s;
}

int colsum_f32(const float* in, int64_t rows, int64_t cols, float* out, float* tmp,
               hipStream_t stream) {
    if (cols <= 0) return AM_OK;
    const int64_t slices = (rows + kColsumSlice - 1) / kColsumSlice;
    const unsigned gx = (unsigned)((cols + 255) / 256);
    if (slices <= 1) {
        hipLaunchKernelGGL(k_colsum, dim3(gx, 1), dim3(256), 0, stream, in, rows, cols,
                           rows > 0 ? rows : 1, out);
        AM_LAUNCH_CHECK();
        return AM_OK;
    }
    hipLaunchKernelGGL(k_colsum, dim3(gx, (unsigned)slices), dim3(256), 0, stream, in, rows, cols,
                       kColsumSlice, tmp);
    AM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_colsum, dim3(gx, 1), dim3(256), 0, stream, tmp, slices, cols, slices,
                       out);
    AM_LAUNCH_CHECK();
    return AM_OK;
}

}  // namespace am

extern "C" const char* am_last_error(void) { return am::last_error(); }

extern "C" size_t am_last_workspace_need(void) { return am::last_workspace_need(); }

extern "C" int am_version(void) { return 10000; /* 1.0.0 */ }

extern "C" int am_timer_enable(int32_t on) {
    am::timer_reset();
    am::g_timer_on = on != 0;
    return AM_OK;
}

extern "C" int am_timer_read(const char* name, double* total_ms, int64_t* count) {
    AM_CHECK_ARG(name && total_ms && count, "NULL argument");
    double tot = 0.0;
    int64_t c = 0;
    for (auto& r : am::g_records) {
        if (r.name != name) continue;
        AM_HIP(hipEventSynchronize(r.b));
        float ms = 0.0f;
        AM_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        tot += ms;
        ++c;
    }
    *total_ms = tot;
    *count = c;
    return AM_OK;
}

namespace am {
static std::mutex g_flag_mu;
static std::vector<std::pair<std::string, int>> g_flags;

int flag_value(const char* name, int dflt) {
    std::lock_guard<std::mutex> lock(g_flag_mu);
    for (const auto& f : g_flags)
        if (f.first == name) return f.second;
    return dflt;
}
}  // namespace am

extern "C" int am_set_flag(const char* name, int32_t value) {
    AM_CHECK_ARG(name, "NULL flag name");
    static const char* known[] = {"table_prepartition", "sdf_bwd_sync_free", "table_atomic"};
    bool ok = false;
    for (const char* k : known) ok = ok || std::string(k) == name;
    AM_CHECK_ARG(ok, "unknown flag %s", name);
    std::lock_guard<std::mutex> lock(am::g_flag_mu);
    for (auto& f : am::g_flags)
        if (f.first == name) {
            f.second = value;
            return AM_OK;
        }
    am::g_flags.emplace_back(name, value);
    return AM_OK;
}

extern "C" int am_timer_units(const char* name, int64_t* units) {
    AM_CHECK_ARG(name && units, "NULL argument");
    *units = 0;
    for (auto& u : am::g_units)
        if (u.name == name) *units = u.units;
    return AM_OK;
}
