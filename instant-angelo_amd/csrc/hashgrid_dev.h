// Device-side multiresolution hash-grid arithmetic (gfx950).
//
// Index and position arithmetic restates tiny-cuda-nn's HashGrid (the
// encoding behind models/network_utils.py:45-47): pos = fmaf(scale, x, 0.5),
// cell = floor(pos), frac = pos - cell; corner c (bit0 x, bit1 y, bit2 z)
// takes cell + bit; dense index x + y*res + z*res^2 or prime hash
// x ^ y*2654435761 ^ z*805459861 (uint32 wrap), then mod the level's rows;
// trilinear weight = product over dims in x,y,z order, blend by fma in corner
// order.  Every step is a single rounded fp32 op (-ffp-contract=off), so the
// indices and positions are bit-exact with the C restatement in the oracle.
#pragma once
#include <hip/hip_fp16.h>
#include "am_common.h"

namespace am {

// One level's parameters (16 bytes: kept in LDS by the fused kernels, where
// lane groups address different levels).
struct LevelParams {
    float scale;
    uint32_t res;
    uint32_t offset;
    uint32_t size_flags;  // rows | bit31 hashed | bit30 rows is a power of two
};

constexpr uint32_t kHashedBit = 1u << 31;
constexpr uint32_t kPow2Bit = 1u << 30;
constexpr uint32_t kSizeMask = kPow2Bit - 1u;

struct GridArgs {
    LevelParams lv[AM_MAX_LEVELS];
    int32_t L;
    int32_t F;
};

GridArgs make_grid_args(const am_grid_desc* g);

// idx mod rows.  Power-of-two levels (every hashed level) mask; a dense
// level's index is below 2 rows for every cell of [0, 1]^3 (coordinates up to
// res, rows >= res^3), so one conditional subtraction does; the division is
// left for inputs outside the unit cube.
__device__ __forceinline__ uint32_t wrap_row(uint32_t idx, uint32_t size_flags) {
    const uint32_t sz = size_flags & kSizeMask;
    if (size_flags & kPow2Bit) return idx & (sz - 1u);
    if (idx < sz) return idx;
    const uint32_t d = idx - sz;
    if (__builtin_expect(d < sz, 1)) return d;
    return idx % sz;
}

__device__ __forceinline__ uint32_t grid_row(const LevelParams& p, uint32_t cx, uint32_t cy,
                                             uint32_t cz) {
    uint32_t idx;
    if (p.size_flags & kHashedBit) {
        idx = cx ^ (cy * 2654435761u) ^ (cz * 805459861u);
    } else {
        idx = cx + cy * p.res + cz * (p.res * p.res);
    }
    return wrap_row(idx, p.size_flags);
}

// The 8 corner rows of `cell` (corner c = cell + bits of c, x = bit 0) with
// the per-axis products formed once: (c+1) * m = c * m + m (mod 2^32).
__device__ __forceinline__ void cell_rows(const LevelParams& lp, const uint32_t cell[3],
                                          uint32_t rows[8]) {
    const bool hashed = (lp.size_flags & kHashedBit) != 0u;
    const uint32_t my = hashed ? 2654435761u : lp.res;
    const uint32_t mz = hashed ? 805459861u : lp.res * lp.res;
    const uint32_t y0 = cell[1] * my, y1 = y0 + my;
    const uint32_t z0 = cell[2] * mz, z1 = z0 + mz;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const uint32_t x = cell[0] + (uint32_t)(c & 1);
        const uint32_t y = (c & 2) ? y1 : y0;
        const uint32_t z = (c & 4) ? z1 : z0;
        const uint32_t idx = hashed ? (x ^ y ^ z) : (x + y + z);
        rows[c] = lp.offset + wrap_row(idx, lp.size_flags);
    }
}

// Cell and fractional position of x at one level.
struct LevelPos {
    uint32_t c[3];
    float f[3];
};

__device__ __forceinline__ LevelPos level_pos(const LevelParams& lp, const float x[3]) {
    LevelPos p;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float v = __fmaf_rn(lp.scale, x[d], 0.5f);
        const float fl = floorf(v);
        p.c[d] = (uint32_t)(int)fl;
        p.f[d] = v - fl;
    }
    return p;
}

// Trilinear weight of corner c (tcnn order: start at 1, multiply x, y, z).
__device__ __forceinline__ float corner_weight(const LevelPos& p, int c) {
    float w = 1.0f;
#pragma unroll
    for (int d = 0; d < 3; ++d) w *= ((c >> d) & 1) ? p.f[d] : (1.0f - p.f[d]);
    return w;
}

// d(weight_c)/d(frac_d).
__device__ __forceinline__ float corner_dweight(const LevelPos& p, int c, int d) {
    float w = ((c >> d) & 1) ? 1.0f : -1.0f;
#pragma unroll
    for (int e = 0; e < 3; ++e)
        if (e != d) w *= ((c >> e) & 1) ? p.f[e] : (1.0f - p.f[e]);
    return w;
}

__device__ __forceinline__ uint32_t corner_row(const LevelParams& lp, const LevelPos& p, int c) {
    return lp.offset + grid_row(lp, p.c[0] + (c & 1), p.c[1] + ((c >> 1) & 1),
                                p.c[2] + ((c >> 2) & 1));
}

// A table row of two features as fp32 (fp32 storage, or fp16 storage of
// the mixed-precision form: the blend stays fp32).
__device__ __forceinline__ float2 load_row2(const float2* __restrict__ t, uint32_t r) { return t[r]; }
__device__ __forceinline__ float2 load_row2(const __half2* __restrict__ t, uint32_t r) {
    return __half22float2(t[r]);
}

// Two-feature level encode (the configuration on the NeuS path).  Optionally
// returns d enc_f / d x_d (including the level scale) for input gradients.
template <bool WITH_DX, class TT>
__device__ __forceinline__ void encode_level_f2(const LevelParams& lp, const float x[3],
                                                const TT* __restrict__ table, float* out2,
                                                float (*dx)[3]) {
    const LevelPos p = level_pos(lp, x);
    uint32_t rows[8];
    cell_rows(lp, p.c, rows);
    float2 v[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) v[c] = load_row2(table, rows[c]);
    float a0 = 0.0f, a1 = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const float w = corner_weight(p, c);
        a0 = __fmaf_rn(w, v[c].x, a0);
        a1 = __fmaf_rn(w, v[c].y, a1);
    }
    out2[0] = a0;
    out2[1] = a1;
    if constexpr (WITH_DX) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const float dw = corner_dweight(p, c, d);
                s0 = __fmaf_rn(dw, v[c].x, s0);
                s1 = __fmaf_rn(dw, v[c].y, s1);
            }
            dx[0][d] = s0 * lp.scale;
            dx[1][d] = s1 * lp.scale;
        }
    }
}

}  // namespace am
