/*
 * ORACLE -- test infrastructure only.  Nothing in the product links, loads or
 * calls this file: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker.
 *
 * Scalar C restatement of the integer / index / position arithmetic on the
 * Instant-angelo NeuS hot path, written so that every result is a fixed
 * sequence of IEEE-754 binary32 operations (built with -ffp-contract=off, an
 * explicit fmaf() wherever the restated algorithm fuses):
 *
 *   - or_level_table     tiny-cuda-nn HashGrid level table (scale, resolution,
 *                        entries, offset, dense/hashed) as used through
 *                        models/network_utils.py:45-47,89-91 (tcnn is not
 *                        vendored in the reference; its published grid rule is
 *                        restated, exp2 taken correctly rounded in fp32).
 *   - or_hashgrid_fwd    per point, per level: fmaf(scale, x, 0.5) -> floor ->
 *                        8 corner indices (dense x+y*res+z*res^2 or the XOR
 *                        prime hash, then mod entries) -> trilinear fma blend,
 *                        level-major output column l*F+f
 *                        (models/network_utils.py:53-58 proves the layout).
 *   - or_ray_aabb        nerfacc 0.3.3 ray_aabb_intersect (slab test, t_min
 *                        clipped to >= 0, 1e10 on a miss), called at
 *                        models/neus.py:209-220 through ray_marching.
 *   - or_ray_marching    nerfacc 0.3.3 ray_marching, AABB contraction, cone
 *                        angle 0: constant step, occupancy test at the
 *                        midpoint, DDA skip over empty cells
 *                        (models/neus.py:209-220).  Build-side additions,
 *                        shared with the HIP kernel: an optional per-ray
 *                        sample cap, and a ray ends when a step no longer
 *                        advances t (the restated loop would spin forever).
 *
 * The nerfacc expression `origin + t_mid * dir` is taken fused (fmaf), the
 * form nvcc's default contraction produces; every other expression there is
 * contraction-invariant (its products are exact).  Unpinnable against
 * NVIDIA's binaries; pinned here by closed forms and invariants (tests/).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define OR_PRIME_Y 2654435761u
#define OR_PRIME_Z 805459861u

int or_level_table(int n_levels, int base_res, float per_level_scale, int log2_T,
                   float* scale, uint32_t* res, uint32_t* offset, uint32_t* size,
                   uint32_t* hashed) {
    if (n_levels <= 0 || n_levels > 32 || base_res <= 0 || log2_T <= 0 || log2_T > 30)
        return -1;
    const float log2_pls = log2f(per_level_scale);
    const uint64_t T = 1ull << log2_T;
    uint64_t off = 0;
    for (int l = 0; l < n_levels; ++l) {
        const float e = (float)l * log2_pls;
        const float p = (float)exp2((double)e);
        const float s = p * (float)base_res - 1.0f;
        const uint32_t r = (uint32_t)ceilf(s) + 1u;
        const double dense = (double)r * (double)r * (double)r;
        uint64_t params = dense > 2147483647.0 ? 2147483647ull
                                               : (uint64_t)r * r * r;
        params = (params + 7ull) / 8ull * 8ull;
        if (params > T) params = T;
        scale[l] = s;
        res[l] = r;
        offset[l] = (uint32_t)off;
        size[l] = (uint32_t)params;
        hashed[l] = dense > (double)params ? 1u : 0u;
        off += params;
    }
    return (int)off;
}

static inline uint32_t or_index(uint32_t is_hashed, uint32_t size, uint32_t res,
                                uint32_t cx, uint32_t cy, uint32_t cz) {
    uint32_t idx;
    if (is_hashed)
        idx = cx ^ (cy * OR_PRIME_Y) ^ (cz * OR_PRIME_Z);
    else
        idx = cx + cy * res + cz * res * res;
    return idx % size;
}

/* Per (point, level): integer cell corner, fractional position, the 8 corner
 * entry indices (level-local) and weights, corner c = bit0:x bit1:y bit2:z. */
void or_hashgrid_corners(const float* x, int64_t n, int L, const float* scale,
                         const uint32_t* res, const uint32_t* size,
                         const uint32_t* hashed, uint32_t* cell /* [n,L,3] */,
                         float* frac /* [n,L,3] */, uint32_t* idx /* [n,L,8] */,
                         float* w /* [n,L,8] */) {
    for (int64_t i = 0; i < n; ++i) {
        for (int l = 0; l < L; ++l) {
            float pos[3];
            uint32_t g[3];
            for (int d = 0; d < 3; ++d) {
                float p = fmaf(scale[l], x[i * 3 + d], 0.5f);
                float fl = floorf(p);
                g[d] = (uint32_t)(int)fl;
                pos[d] = p - fl;
            }
            const int64_t o = i * L + l;
            if (cell) for (int d = 0; d < 3; ++d) cell[o * 3 + d] = g[d];
            if (frac) for (int d = 0; d < 3; ++d) frac[o * 3 + d] = pos[d];
            for (int c = 0; c < 8; ++c) {
                float wt = 1.0f;
                uint32_t lc[3];
                for (int d = 0; d < 3; ++d) {
                    if (c & (1 << d)) { wt *= pos[d]; lc[d] = g[d] + 1u; }
                    else { wt *= 1.0f - pos[d]; lc[d] = g[d]; }
                }
                if (idx) idx[o * 8 + c] = or_index(hashed[l], size[l], res[l], lc[0], lc[1], lc[2]);
                if (w) w[o * 8 + c] = wt;
            }
        }
    }
}

/* Forward encode, fp32 table [entries, F], output [n, L*F] level-major. */
void or_hashgrid_fwd(const float* x, int64_t n, const float* table, int L, int F,
                     const float* scale, const uint32_t* res, const uint32_t* offset,
                     const uint32_t* size, const uint32_t* hashed, float* out) {
    for (int64_t i = 0; i < n; ++i) {
        for (int l = 0; l < L; ++l) {
            float pos[3];
            uint32_t g[3];
            for (int d = 0; d < 3; ++d) {
                float p = fmaf(scale[l], x[i * 3 + d], 0.5f);
                float fl = floorf(p);
                g[d] = (uint32_t)(int)fl;
                pos[d] = p - fl;
            }
            float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int c = 0; c < 8; ++c) {
                float wt = 1.0f;
                uint32_t lc[3];
                for (int d = 0; d < 3; ++d) {
                    if (c & (1 << d)) { wt *= pos[d]; lc[d] = g[d] + 1u; }
                    else { wt *= 1.0f - pos[d]; lc[d] = g[d]; }
                }
                const uint32_t e = offset[l] + or_index(hashed[l], size[l], res[l], lc[0], lc[1], lc[2]);
                for (int f = 0; f < F; ++f)
                    acc[f] = fmaf(wt, table[(int64_t)e * F + f], acc[f]);
            }
            for (int f = 0; f < F; ++f) out[i * (int64_t)L * F + l * F + f] = acc[f];
        }
    }
}

/* ---- nerfacc 0.3.3 ray_aabb_intersect ---- */
static void or_aabb_one(const float* o, const float* d, const float* aabb,
                        float* t_near, float* t_far) {
    float tmin = (aabb[0] - o[0]) / d[0];
    float tmax = (aabb[3] - o[0]) / d[0];
    if (tmin > tmax) { float t = tmin; tmin = tmax; tmax = t; }
    float tymin = (aabb[1] - o[1]) / d[1];
    float tymax = (aabb[4] - o[1]) / d[1];
    if (tymin > tymax) { float t = tymin; tymin = tymax; tymax = t; }
    if (tmin > tymax || tymin > tmax) { *t_near = 1e10f; *t_far = 1e10f; return; }
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    float tzmin = (aabb[2] - o[2]) / d[2];
    float tzmax = (aabb[5] - o[2]) / d[2];
    if (tzmin > tzmax) { float t = tzmin; tzmin = tzmax; tzmax = t; }
    if (tmin > tzmax || tzmin > tmax) { *t_near = 1e10f; *t_far = 1e10f; return; }
    if (tzmin > tmin) tmin = tzmin;
    if (tzmax < tmax) tmax = tzmax;
    *t_near = tmin > 0.0f ? tmin : 0.0f;
    *t_far = tmax;
}

void or_ray_aabb(const float* rays_o, const float* rays_d, int64_t n, const float* aabb,
                 float* t_near, float* t_far) {
    for (int64_t i = 0; i < n; ++i)
        or_aabb_one(rays_o + 3 * i, rays_d + 3 * i, aabb, t_near + i, t_far + i);
}

/* ---- nerfacc 0.3.3 ray_marching (AABB contraction, cone_angle = 0) ---- */
static int or_occupied(const float xyz[3], const float* roi, int R, const uint8_t* grid) {
    for (int d = 0; d < 3; ++d)
        if (xyz[d] < roi[d] || xyz[d] > roi[3 + d]) return 0;
    int ix[3];
    for (int d = 0; d < 3; ++d) {
        float u = (xyz[d] - roi[d]) / (roi[3 + d] - roi[d]);
        int v = (int)(u * (float)R);
        if (v < 0) v = 0;
        if (v > R - 1) v = R - 1;
        ix[d] = v;
    }
    return grid[(int64_t)ix[0] * R * R + (int64_t)ix[1] * R + ix[2]] != 0;
}

static float or_dist_next_voxel(const float xyz[3], const float dir[3], const float inv[3],
                                const float* roi, int R) {
    float t = INFINITY;
    for (int d = 0; d < 3; ++d) {
        const float range = roi[3 + d] - roi[d];
        const float u = ((xyz[d] - roi[d]) / range) * (float)R;
        const float sg = copysignf(1.0f, dir[d]);
        const float tt = ((floorf(u + 0.5f + 0.5f * sg) - u) * inv[d]) / (float)R * range;
        t = fminf(t, tt);
    }
    return fmaxf(t, 0.0f);
}

/* March one ray; when out != NULL writes up to `cap` samples.  Returns count. */
static int or_march_one(const float* o, const float* d, float near, float far,
                        const float* roi, int R, const uint8_t* grid, float step,
                        int cap, float* t_starts, float* t_ends) {
    const float inv[3] = {1.0f / d[0], 1.0f / d[1], 1.0f / d[2]};
    int j = 0;
    float t0 = near;
    float t1 = t0 + step;
    float t_mid = (t0 + t1) * 0.5f;
    while (t_mid < far) {
        if (cap > 0 && j >= cap) break;
        float xyz[3];
        for (int k = 0; k < 3; ++k) xyz[k] = fmaf(t_mid, d[k], o[k]);
        if (or_occupied(xyz, roi, R, grid)) {
            if (t_starts) { t_starts[j] = t0; t_ends[j] = t1; }
            ++j;
            const float prev = t_mid;
            t0 = t1;
            t1 = t0 + step;
            t_mid = (t0 + t1) * 0.5f;
            if (!(t_mid > prev)) break;
        } else {
            const float target = t_mid + or_dist_next_voxel(xyz, d, inv, roi, R);
            float tt = t_mid;
            int stuck = 0;
            do {
                const float nt = tt + step;
                if (!(nt > tt)) { stuck = 1; break; }
                tt = nt;
            } while (tt < target);
            if (stuck) break;
            t_mid = tt;
            t0 = t_mid - step * 0.5f;
            t1 = t_mid + step * 0.5f;
        }
    }
    return j;
}

/* Pass 1: per-ray counts.  Returns total. */
int64_t or_ray_marching_count(const float* rays_o, const float* rays_d, const float* t_near,
                              const float* t_far, int64_t n, const float* roi, int R,
                              const uint8_t* grid, float step, int cap, int32_t* counts) {
    int64_t total = 0;
    for (int64_t i = 0; i < n; ++i) {
        counts[i] = or_march_one(rays_o + 3 * i, rays_d + 3 * i, t_near[i], t_far[i], roi, R,
                                 grid, step, cap, NULL, NULL);
        total += counts[i];
    }
    return total;
}

/* Pass 2: packed, ray-sorted samples (caller provides total-sized outputs). */
void or_ray_marching_fill(const float* rays_o, const float* rays_d, const float* t_near,
                          const float* t_far, int64_t n, const float* roi, int R,
                          const uint8_t* grid, float step, int cap, int32_t* ray_indices,
                          float* t_starts, float* t_ends) {
    int64_t base = 0;
    for (int64_t i = 0; i < n; ++i) {
        int c = or_march_one(rays_o + 3 * i, rays_d + 3 * i, t_near[i], t_far[i], roi, R, grid,
                             step, cap, t_starts + base, t_ends + base);
        for (int k = 0; k < c; ++k) ray_indices[base + k] = (int32_t)i;
        base += c;
    }
}

/* ---- background marching: nerfacc 0.3.3 ray_marching with
 * ContractionType.UN_BOUNDED_SPHERE and cone_angle > 0, as called by
 * models/neus.py:141-171 (library absent from the reference tree; restated
 * from its published kernel: contracted scenes step t0 -> t1 = t0 +
 * clamp(t0 * cone, step, 1e10) and keep a sample when the contracted
 * midpoint's cell is occupied; no voxel skip). */
static int or_occupied_unbounded(const float xyz[3], const float* roi, int R,
                                 const uint8_t* grid) {
    if (!grid) return 1;
    float x[3];
    for (int d = 0; d < 3; ++d) x[d] = ((xyz[d] - roi[d]) / (roi[3 + d] - roi[d])) * 2.0f - 1.0f;
    const float mag = sqrtf((x[0] * x[0] + x[1] * x[1]) + x[2] * x[2]);
    if (mag > 1.0f) {
        const float k = 2.0f - 1.0f / mag;
        for (int d = 0; d < 3; ++d) x[d] = k * (x[d] / mag);
    }
    int ix[3];
    for (int d = 0; d < 3; ++d) {
        const float u = x[d] * 0.25f + 0.5f;
        int v = (int)(u * (float)R);
        if (v < 0) v = 0;
        if (v > R - 1) v = R - 1;
        ix[d] = v;
    }
    return grid[(int64_t)ix[0] * R * R + (int64_t)ix[1] * R + ix[2]] != 0;
}

static float or_cone_dt(float t, float cone, float step) {
    return fminf(fmaxf(t * cone, step), 1e10f);
}

static int or_march_one_bg(const float* o, const float* d, float near, float far,
                           const float* roi, int R, const uint8_t* grid, float step, float cone,
                           int cap, float* t_starts, float* t_ends) {
    int j = 0;
    float t0 = near;
    float t1 = t0 + or_cone_dt(t0, cone, step);
    float t_mid = (t0 + t1) * 0.5f;
    while (t_mid < far) {
        if (cap > 0 && j >= cap) break;
        float xyz[3];
        for (int k = 0; k < 3; ++k) xyz[k] = fmaf(t_mid, d[k], o[k]);
        if (or_occupied_unbounded(xyz, roi, R, grid)) {
            if (t_starts) { t_starts[j] = t0; t_ends[j] = t1; }
            ++j;
        }
        const float prev = t_mid;
        t0 = t1;
        t1 = t0 + or_cone_dt(t0, cone, step);
        t_mid = (t0 + t1) * 0.5f;
        if (!(t_mid > prev)) break;
    }
    return j;
}

int64_t or_march_bg_count(const float* rays_o, const float* rays_d, const float* t_near,
                          const float* t_far, int64_t n, const float* roi, int R,
                          const uint8_t* grid, float step, float cone, int cap, int32_t* counts) {
    int64_t total = 0;
    for (int64_t i = 0; i < n; ++i) {
        counts[i] = or_march_one_bg(rays_o + 3 * i, rays_d + 3 * i, t_near[i], t_far[i], roi, R,
                                    grid, step, cone, cap, NULL, NULL);
        total += counts[i];
    }
    return total;
}

void or_march_bg_fill(const float* rays_o, const float* rays_d, const float* t_near,
                      const float* t_far, int64_t n, const float* roi, int R, const uint8_t* grid,
                      float step, float cone, int cap, int32_t* ray_indices, float* t_starts,
                      float* t_ends) {
    int64_t base = 0;
    for (int64_t i = 0; i < n; ++i) {
        int c = or_march_one_bg(rays_o + 3 * i, rays_d + 3 * i, t_near[i], t_far[i], roi, R, grid,
                                step, cone, cap, t_starts + base, t_ends + base);
        for (int k = 0; k < c; ++k) ray_indices[base + k] = (int32_t)i;
        base += c;
    }
}

/* Early-stop visibility of nerfacc 0.3.3 ray_marching with a sigma_fn
 * (alpha_thre 0, early_stop_eps): alpha = 1 - exp(-sigma * (t1 - t0)),
 * transmittance T_j = prod_{i<j} (1 - alpha_i) in sample order, a sample kept
 * while T_j >= eps.  T never grows along a ray, so the kept samples are a
 * prefix; writes each ray's prefix length. */
void or_visible_prefix(const float* sigma, const float* t_starts, const float* t_ends,
                       const int32_t* packed, int64_t n_rays, float eps, int32_t* keep) {
    for (int64_t r = 0; r < n_rays; ++r) {
        const int32_t b = packed[2 * r], c = packed[2 * r + 1];
        float T = 1.0f;
        int32_t k = 0;
        for (; k < c; ++k) {
            if (!(T >= eps)) break;
            const float alpha = 1.0f - expf(-(sigma[b + k] * (t_ends[b + k] - t_starts[b + k])));
            T = T * (1.0f - alpha);
        }
        keep[r] = k;
    }
}
