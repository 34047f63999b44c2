// Error reporting and version for the C-ABI.
#include <cstdarg>
#include <cstdio>
#include "am_common.h"

namespace am {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

const char* last_error() { return g_err; }

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

extern "C" int am_version(void) { return 10000; /* 1.0.0 */ }
