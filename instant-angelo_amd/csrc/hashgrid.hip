// Multiresolution hash grid: level table, encode forward, deterministic
// backward (C-ABI half of ProgressiveBandHashGrid / tcnn HashGrid,
// models/network_utils.py:40-66,83-93).
#include <cmath>
#include "hashgrid_dev.h"
#include "table_grad.h"

namespace am {

GridArgs make_grid_args(const am_grid_desc* g) {
    GridArgs a{};
    for (int l = 0; l < g->n_levels; ++l) {
        LevelParams& p = a.lv[l];
        p.scale = g->scale[l];
        p.res = g->res[l];
        p.offset = g->offset[l];
        p.size_flags = g->size[l];
        if ((g->size[l] & (g->size[l] - 1u)) == 0u) p.size_flags |= kPow2Bit;
        if ((g->hashed_mask >> l) & 1u) p.size_flags |= kHashedBit;
    }
    a.L = g->n_levels;
    a.F = g->n_features;
    return a;
}

static int check_grid(const am_grid_desc* g) {
    AM_CHECK_ARG(g != nullptr, "grid descriptor is NULL");
    AM_CHECK_ARG(g->n_levels >= 1 && g->n_levels <= AM_MAX_LEVELS, "n_levels %d out of [1,32]",
                 g->n_levels);
    AM_CHECK_ARG(g->n_features == 1 || g->n_features == 2 || g->n_features == 4 ||
                     g->n_features == 8,
                 "n_features %d not in {1,2,4,8}", g->n_features);
    AM_CHECK_ARG(g->n_entries > 0, "empty table");
    return AM_OK;
}

template <int F>
struct FVec {
    float v[F];
};

template <int F>
__device__ __forceinline__ FVec<F> load_row(const float* __restrict__ t, uint32_t row) {
    FVec<F> r;
    if constexpr (F == 1) {
        r.v[0] = t[row];
    } else if constexpr (F == 2) {
        const float2 a = reinterpret_cast<const float2*>(t)[row];
        r.v[0] = a.x;
        r.v[1] = a.y;
    } else {
#pragma unroll
        for (int q = 0; q < F / 4; ++q) {
            const float4 a = reinterpret_cast<const float4*>(t)[(size_t)row * (F / 4) + q];
            r.v[4 * q + 0] = a.x;
            r.v[4 * q + 1] = a.y;
            r.v[4 * q + 2] = a.z;
            r.v[4 * q + 3] = a.w;
        }
    }
    return r;
}

template <int F>
__global__ void __launch_bounds__(256) k_hg_fwd(GridArgs g, const float* __restrict__ x, int64_t n,
                                                const float* __restrict__ table, int active,
                                                float* __restrict__ out,
                                                const int32_t* __restrict__ order) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t i = order ? (int64_t)order[k] : k;   // the caller's row
    const float xi[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    float* o = out + i * (int64_t)g.L * F;
    for (int l = 0; l < g.L; ++l) {
        float acc[F];
#pragma unroll
        for (int f = 0; f < F; ++f) acc[f] = 0.0f;
        if (l < active) {
            const LevelParams lp = g.lv[l];
            const LevelPos p = level_pos(lp, xi);
            FVec<F> v[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) v[c] = load_row<F>(table, corner_row(lp, p, c));
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const float w = corner_weight(p, c);
#pragma unroll
                for (int f = 0; f < F; ++f) acc[f] = __fmaf_rn(w, v[c].v[f], acc[f]);
            }
        }
#pragma unroll
        for (int f = 0; f < F; ++f) o[l * F + f] = acc[f];
    }
}

// One thread per (level < active, point); pair slot ((l*n + i)*8 + c).
template <int F>
__global__ void __launch_bounds__(256) k_hg_emit(GridArgs g, const float* __restrict__ x, int64_t n,
                                                 const float* __restrict__ dy, int active,
                                                 uint32_t* __restrict__ keys,
                                                 Vals<F>* __restrict__ vals,
                                                 const int32_t* __restrict__ order) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * active) return;
    const int l = (int)(t / n);
    const int64_t k = t - (int64_t)l * n;
    const int64_t i = order ? (int64_t)order[k] : k;   // the caller's row
    const float xi[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    const LevelParams lp = g.lv[l];
    const LevelPos p = level_pos(lp, xi);
    float gy[F];
#pragma unroll
    for (int f = 0; f < F; ++f) gy[f] = dy[i * (int64_t)g.L * F + l * F + f];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const float w = corner_weight(p, c);
        Vals<F> v;
#pragma unroll
        for (int f = 0; f < F; ++f) v.v[f] = w * gy[f];
        keys[t * 8 + c] = corner_row(lp, p, c);
        vals[t * 8 + c] = v;
    }
}

// dL/dx = sum_l sum_f dy[l,f] * d enc[l,f] / dx.
template <int F>
__global__ void __launch_bounds__(256) k_hg_dx(GridArgs g, const float* __restrict__ x, int64_t n,
                                               const float* __restrict__ table,
                                               const float* __restrict__ dy, int active,
                                               float* __restrict__ dx,
                                               const int32_t* __restrict__ order) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t i = order ? (int64_t)order[k] : k;   // the caller's row
    const float xi[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    float acc[3] = {0.0f, 0.0f, 0.0f};
    for (int l = 0; l < active; ++l) {
        const LevelParams lp = g.lv[l];
        const LevelPos p = level_pos(lp, xi);
        FVec<F> v[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) v[c] = load_row<F>(table, corner_row(lp, p, c));
        float gy[F];
#pragma unroll
        for (int f = 0; f < F; ++f) gy[f] = dy[i * (int64_t)g.L * F + l * F + f];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            float s = 0.0f;
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                float dv = 0.0f;
#pragma unroll
                for (int f = 0; f < F; ++f) dv = __fmaf_rn(gy[f], v[c].v[f], dv);
                s = __fmaf_rn(corner_dweight(p, c, d), dv, s);
            }
            acc[d] = __fmaf_rn(s, lp.scale, acc[d]);
        }
    }
    dx[3 * i + 0] = acc[0];
    dx[3 * i + 1] = acc[1];
    dx[3 * i + 2] = acc[2];
}

template <int F>
static int hg_fwd(const GridArgs& ga, const float* x, int64_t n, const float* table, int active,
                  float* out, const int32_t* order, hipStream_t s) {
    {
        KernelTimer kt_("hg_fwd", s);
        hipLaunchKernelGGL(k_hg_fwd<F>, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, ga, x, n,
                           table, active, out, order);
    }
    AM_LAUNCH_CHECK();
    return AM_OK;
}

template <int F>
static int hg_bwd(const GridArgs& ga, uint32_t n_entries, const float* x, int64_t n,
                  const float* table, const float* dy, int active, float* dtable, float* dx,
                  void* ws, size_t ws_bytes, const int32_t* order, hipStream_t s) {
    if (dx) {
        hipLaunchKernelGGL(k_hg_dx<F>, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, ga, x,
                           n, table, dy, active, dx, order);
        AM_LAUNCH_CHECK();
    }
    if (dtable) {
        const int64_t m = n * active * 8;
        if (m == 0) return fill_f32(dtable, (int64_t)n_entries * F, 0.0f, s);
        WsCarver c(ws, ws_bytes);
        TableReduceWs w;
        int rc = table_reduce_carve(c, F, m, n_entries, &w);
        if (rc) return rc;
        if (!c.ok()) return workspace_short(c.used, ws_bytes);
        {
            KernelTimer kt_("hg_emit", s);
            hipLaunchKernelGGL(k_hg_emit<F>, dim3((unsigned)ceil_div(n * active, 256)), dim3(256), 0,
                               s, ga, x, n, dy, active, w.keys[0],
                               reinterpret_cast<Vals<F>*>(w.vals[0]), order);
        }
        AM_LAUNCH_CHECK();
        timer_units("encode_table_pairs", w.m);
        return table_reduce_run(w, F, n_entries, dtable, s);
    }
    return AM_OK;
}

}  // namespace am

using namespace am;

extern "C" int am_grid_init(am_grid_desc* g, int32_t n_levels, int32_t n_features,
                            int32_t log2_hashmap_size, int32_t base_resolution,
                            float per_level_scale, const uint32_t* res_override) {
    AM_CHECK_ARG(g != nullptr, "grid descriptor is NULL");
    AM_CHECK_ARG(n_levels >= 1 && n_levels <= AM_MAX_LEVELS, "n_levels %d out of [1,32]", n_levels);
    AM_CHECK_ARG(n_features == 1 || n_features == 2 || n_features == 4 || n_features == 8,
                 "n_features %d not in {1,2,4,8}", n_features);
    AM_CHECK_ARG(log2_hashmap_size >= 1 && log2_hashmap_size <= 30, "log2_hashmap_size %d",
                 log2_hashmap_size);
    AM_CHECK_ARG(base_resolution >= 1, "base_resolution %d", base_resolution);
    AM_CHECK_ARG(per_level_scale >= 1.0f, "per_level_scale %g < 1", (double)per_level_scale);
    am_grid_desc d{};
    d.n_levels = n_levels;
    d.n_features = n_features;
    d.log2_hashmap_size = log2_hashmap_size;
    d.base_resolution = base_resolution;
    d.per_level_scale = per_level_scale;
    const float log2_pls = std::log2(per_level_scale);
    const uint64_t T = 1ull << log2_hashmap_size;
    uint64_t off = 0;
    for (int l = 0; l < n_levels; ++l) {
        float s;
        uint32_t r;
        if (res_override) {
            r = res_override[l];
            AM_CHECK_ARG(r >= 2, "res_override[%d] = %u < 2", l, r);
            s = (float)(r - 1u);
        } else {
            const float e = (float)l * log2_pls;
            const float p = (float)std::exp2((double)e);  // correctly rounded fp32 exp2
            s = p * (float)base_resolution;
            s = s - 1.0f;
            r = (uint32_t)std::ceil(s) + 1u;
        }
        const double dense = (double)r * (double)r * (double)r;
        uint64_t params = dense > 2147483647.0 ? 2147483647ull : (uint64_t)r * r * r;
        params = (params + 7ull) / 8ull * 8ull;
        if (params > T) params = T;
        d.scale[l] = s;
        d.res[l] = r;
        d.offset[l] = (uint32_t)off;
        d.size[l] = (uint32_t)params;
        if (dense > (double)params) d.hashed_mask |= 1u << l;
        off += params;
        AM_CHECK_ARG(off < (1ull << 31), "table exceeds 2^31 rows");
    }
    d.n_entries = (uint32_t)off;
    *g = d;
    return AM_OK;
}

static int hashgrid_fwd(const am_grid_desc* g, const float* x, int64_t n, const float* table,
                        int32_t active_levels, float* out, const int32_t* order, void* stream) {
    if (int rc = check_grid(g)) return rc;
    AM_CHECK_ARG(n >= 0, "n < 0");
    if (n == 0) return AM_OK;
    AM_CHECK_ARG(x && table && out, "NULL buffer");
    AM_CHECK_ARG(active_levels >= 0 && active_levels <= g->n_levels, "active_levels %d",
                 active_levels);
    const GridArgs ga = make_grid_args(g);
    hipStream_t s = to_stream(stream);
    switch (g->n_features) {
        case 1: return hg_fwd<1>(ga, x, n, table, active_levels, out, order, s);
        case 2: return hg_fwd<2>(ga, x, n, table, active_levels, out, order, s);
        case 4: return hg_fwd<4>(ga, x, n, table, active_levels, out, order, s);
        default: return hg_fwd<8>(ga, x, n, table, active_levels, out, order, s);
    }
}

extern "C" int am_hashgrid_fwd(const am_grid_desc* g, const float* x, int64_t n,
                               const float* table, int32_t active_levels, float* out,
                               void* stream) {
    return hashgrid_fwd(g, x, n, table, active_levels, out, nullptr, stream);
}

extern "C" int am_hashgrid_fwd_ordered(const am_grid_desc* g, const int32_t* order,
                                       const float* x, int64_t n, const float* table,
                                       int32_t active_levels, float* out, void* stream) {
    AM_CHECK_ARG(order || n == 0, "NULL order");
    return hashgrid_fwd(g, x, n, table, active_levels, out, order, stream);
}

extern "C" size_t am_hashgrid_bwd_workspace_size(const am_grid_desc* g, int64_t n,
                                                 int32_t active_levels) {
    if (check_grid(g) || n < 0 || active_levels < 0 || active_levels > g->n_levels) return 0;
    WsCarver c(nullptr, 0);
    TableReduceWs w;
    if (table_reduce_carve(c, g->n_features, n * active_levels * 8, g->n_entries, &w)) return 0;
    return c.used + 256;
}

static int hashgrid_bwd(const am_grid_desc* g, const float* x, int64_t n, const float* table,
                        const float* dy, int32_t active_levels, float* dtable, float* dx,
                        void* workspace, size_t workspace_bytes, const int32_t* order,
                        void* stream) {
    if (int rc = check_grid(g)) return rc;
    AM_CHECK_ARG(n >= 0, "n < 0");
    AM_CHECK_ARG(active_levels >= 0 && active_levels <= g->n_levels, "active_levels %d",
                 active_levels);
    AM_CHECK_ARG(x || n == 0, "x is NULL");
    AM_CHECK_ARG(dy || n == 0, "dy is NULL");
    AM_CHECK_ARG(!dx || table, "dx needs the table");
    AM_CHECK_ARG(!dtable || workspace || n == 0 || active_levels == 0, "workspace is NULL");
    const GridArgs ga = make_grid_args(g);
    hipStream_t s = to_stream(stream);
    switch (g->n_features) {
        case 1: return hg_bwd<1>(ga, g->n_entries, x, n, table, dy, active_levels, dtable, dx, workspace, workspace_bytes, order, s);
        case 2: return hg_bwd<2>(ga, g->n_entries, x, n, table, dy, active_levels, dtable, dx, workspace, workspace_bytes, order, s);
        case 4: return hg_bwd<4>(ga, g->n_entries, x, n, table, dy, active_levels, dtable, dx, workspace, workspace_bytes, order, s);
        default: return hg_bwd<8>(ga, g->n_entries, x, n, table, dy, active_levels, dtable, dx, workspace, workspace_bytes, order, s);
    }
}

extern "C" int am_hashgrid_bwd(const am_grid_desc* g, const float* x, int64_t n,
                               const float* table, const float* dy, int32_t active_levels,
                               float* dtable, float* dx, void* workspace, size_t workspace_bytes,
                               void* stream) {
    return hashgrid_bwd(g, x, n, table, dy, active_levels, dtable, dx, workspace,
                        workspace_bytes, nullptr, stream);
}

extern "C" int am_hashgrid_bwd_ordered(const am_grid_desc* g, const int32_t* order,
                                       const float* x, int64_t n, const float* table,
                                       const float* dy, int32_t active_levels, float* dtable,
                                       float* dx, void* workspace, size_t workspace_bytes,
                                       void* stream) {
    AM_CHECK_ARG(order || n == 0, "NULL order");
    return hashgrid_bwd(g, x, n, table, dy, active_levels, dtable, dx, workspace,
                        workspace_bytes, order, stream);
}
