// Occupancy-grid update (nerfacc 0.3.3 OccupancyGrid._update through
// every_n_step, models/neus.py:63-68,90-111): jittered cell points -> world
// (AABB contract_inv) -> SDF (base evaluation of the fused geometry kernel)
// -> NeuS alpha estimate of occ_eval_fn (models/neus.py:90-101) -> EMA max
// update occs = max(occs*decay, occ) -> binary = occs > min(mean, thre).
//
// Cells listed more than once take the max of their estimates (the
// scatter-max form nerfacc notes as equivalent; deterministic here), and
// the decay is applied once per listed cell.
#include "am_common.h"

extern "C" int am_sdf_fwd(const am_grid_desc* g, const am_mlp_desc* m, const am_sdf_desc* sd,
                          const float* pts, const float* randn, int64_t n, const void* table,
                          const float* params, float* sdf, float* grad, float* normal,
                          float* feature, float* laplace, float* fd_values, float* enc_store,
                          void* stream);

namespace am {

__global__ void k_occ_points(const int32_t* __restrict__ cells, const float* __restrict__ jitter,
                             int64_t m, int res, float r, float* __restrict__ pts) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int64_t c = cells ? (int64_t)cells[i] : i;
    const int64_t cx = c / ((int64_t)res * res), cy = (c / res) % res, cz = c % res;
    const int64_t cc[3] = {cx, cy, cz};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        // x = (coord + u) / res in [0,1]; world = x * (2r) + (-r)
        const float u01 = ((float)cc[d] + jitter[3 * i + d]) / (float)res;
        pts[3 * i + d] = u01 * (2.0f * r) + (-r);
    }
}

__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// alpha estimate of occ_eval_fn, scattered as the max per cell (float bits of
// a non-negative value order like ints; -1 marks untouched cells).
__global__ void k_occ_alpha(const int32_t* __restrict__ cells, const float* __restrict__ sdf,
                            int64_t m, const float* __restrict__ inv_s_p, float half_step,
                            int32_t* __restrict__ best) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    float s = inv_s_p[0];
    s = fminf(fmaxf(s, 1e-6f), 1e6f);
    const float f = sdf[i];
    const float nxt = f - half_step, prv = f + half_step;
    const float pc = 1.0f / (1.0f + expf(-(prv * s)));
    const float nc = 1.0f / (1.0f + expf(-(nxt * s)));
    float a = ((pc - nc) + 1e-5f) / (pc + 1e-5f);
    a = fminf(fmaxf(a, 0.0f), 1.0f);
    const int64_t c = cells ? (int64_t)cells[i] : i;
    atomicMax(&best[c], __float_as_int(a));
}

__global__ void k_occ_ema(int64_t n_cells, const int32_t* __restrict__ best, float decay,
                          float* __restrict__ occs, float* __restrict__ block_sums) {
    __shared__ float red[256];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float v = 0.0f;
    if (i < n_cells) {
        v = occs[i];
        const int32_t b = best[i];
        if (b >= 0) {
            v = fmaxf(v * decay, __int_as_float(b));
            occs[i] = v;
        }
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) block_sums[blockIdx.x] = red[0];
}

__global__ void k_occ_binary(int64_t n_cells, const float* __restrict__ occs,
                             const float* __restrict__ total, float thre,
                             uint8_t* __restrict__ binary) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_cells) return;
    const float mean = total[0] / (float)n_cells;
    const float t = fminf(mean, thre);
    binary[i] = occs[i] > t ? 1 : 0;
}

}  // namespace am

using namespace am;

extern "C" size_t am_occ_update_workspace_size(int32_t grid_res, int64_t m) {
    if (grid_res < 1 || m < 0) return 0;
    const int64_t cells = (int64_t)grid_res * grid_res * grid_res;
    const int64_t blocks = ceil_div(cells, 256);
    WsCarver c(nullptr, 0);
    c.take<float>((size_t)(3 * m + 3));
    c.take<float>((size_t)m + 1);
    c.take<int32_t>((size_t)cells);
    c.take<float>((size_t)blocks);
    c.take<float>((size_t)colsum_tmp_floats(blocks, 1) + 1);
    c.take<float>(1);
    return c.used + 256;
}

extern "C" int am_occ_update(const am_grid_desc* g, const am_mlp_desc* mlp, const am_sdf_desc* s,
                             const void* table, const float* params, const float* inv_s,
                             float step, const int32_t* cells, const float* jitter, int64_t m,
                             int32_t grid_res, float ema_decay, float occ_thre, float* occs,
                             uint8_t* binary, void* workspace, size_t workspace_bytes,
                             void* stream) {
    AM_CHECK_ARG(g && mlp && s, "NULL descriptor");
    AM_CHECK_ARG(grid_res >= 1 && grid_res <= 1024, "grid_res %d", grid_res);
    AM_CHECK_ARG(m >= 0, "m < 0");
    AM_CHECK_ARG(table && params && inv_s && occs && binary && workspace, "NULL buffer");
    AM_CHECK_ARG(m == 0 || jitter, "NULL jitter");
    const int64_t n_cells = (int64_t)grid_res * grid_res * grid_res;
    AM_CHECK_ARG(cells || m == n_cells, "cells == NULL needs m == grid_res^3");
    hipStream_t st = to_stream(stream);
    WsCarver c(workspace, workspace_bytes);
    float* pts = c.take<float>((size_t)(3 * m + 3));
    float* sdf = c.take<float>((size_t)m + 1);
    int32_t* best = c.take<int32_t>((size_t)n_cells);
    const int64_t blocks = ceil_div(n_cells, 256);
    float* bsum = c.take<float>((size_t)blocks);
    float* ctmp = c.take<float>((size_t)colsum_tmp_floats(blocks, 1) + 1);
    float* total = c.take<float>(1);
    if (!c.ok()) return workspace_short(c.used, workspace_bytes);
    am_sdf_desc sd = *s;
    sd.with_grad = 0;
    sd.with_laplace = 0;
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)(blocks < 8192 ? blocks : 8192)), dim3(256), 0, st,
                       best, n_cells, -1);
    AM_LAUNCH_CHECK();
    if (m > 0) {
        hipLaunchKernelGGL(k_occ_points, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, st, cells,
                           jitter, m, grid_res, s->radius, pts);
        AM_LAUNCH_CHECK();
        TimerAlias alias("occ_sdf");
        if (int rc = am_sdf_fwd(g, mlp, &sd, pts, nullptr, m, table, params, sdf, nullptr, nullptr,
                                nullptr, nullptr, nullptr, nullptr, stream))
            return rc;
        hipLaunchKernelGGL(k_occ_alpha, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, st, cells,
                           sdf, m, inv_s, step * 0.5f, best);
        AM_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_occ_ema, dim3((unsigned)blocks), dim3(256), 0, st, n_cells, best, ema_decay,
                       occs, bsum);
    AM_LAUNCH_CHECK();
    if (int rc = colsum_f32(bsum, blocks, 1, total, ctmp, st)) return rc;
    hipLaunchKernelGGL(k_occ_binary, dim3((unsigned)blocks), dim3(256), 0, st, n_cells, occs, total,
                       occ_thre, binary);
    AM_LAUNCH_CHECK();
    return AM_OK;
}
