// Ray-AABB intersection and occupancy-grid ray marching (nerfacc 0.3.3
// ray_aabb_intersect + ray_marching with AABB contraction and cone angle 0,
// as called at models/neus.py:209-220): count pass -> exclusive scan ->
// fill pass, giving packed ray-sorted (ray_indices, t_starts, t_ends).
//
// The arithmetic is the same IEEE sequence as the C restatement in the
// oracle (built with -ffp-contract=off; `origin + t_mid*dir` fused), so
// sample intervals and ray indices are bit-exact with it.  Build-side
// additions shared with the oracle: an optional per-ray sample cap and the
// end of a ray when a step no longer advances t.
#include <cmath>
#include <rocprim/device/device_scan.hpp>
#include "am_common.h"

namespace am {

struct MarchArgs {
    float roi[6];
    int res;
    float step;
    int cap;
    int has_grid;
    // Division by a power of two is an exact scaling, so x / range and
    // x * (1 / range) are the same IEEE result: the divisions of the marching
    // rule become multiplications where the roi extent (and the grid
    // resolution) is a power of two, as for the usual [-r, r]^3 with r = 2^k.
    float inv_range[3];
    int range_pow2;  // bit d: roi extent of axis d is a power of two
    float inv_res;
    int res_pow2;
};

__device__ __forceinline__ float div_range(const MarchArgs& a, float x, int d) {
    return ((a.range_pow2 >> d) & 1) ? x * a.inv_range[d] : x / (a.roi[3 + d] - a.roi[d]);
}

__device__ __forceinline__ float div_res(const MarchArgs& a, float x) {
    return a.res_pow2 ? x * a.inv_res : x / (float)a.res;
}

__device__ __forceinline__ void aabb_one(const float* o, const float* d, const float* aabb,
                                         float* t_near, float* t_far) {
    float tmin = (aabb[0] - o[0]) / d[0];
    float tmax = (aabb[3] - o[0]) / d[0];
    if (tmin > tmax) { const float t = tmin; tmin = tmax; tmax = t; }
    float tymin = (aabb[1] - o[1]) / d[1];
    float tymax = (aabb[4] - o[1]) / d[1];
    if (tymin > tymax) { const float t = tymin; tymin = tymax; tymax = t; }
    if (tmin > tymax || tymin > tmax) { *t_near = 1e10f; *t_far = 1e10f; return; }
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    float tzmin = (aabb[2] - o[2]) / d[2];
    float tzmax = (aabb[5] - o[2]) / d[2];
    if (tzmin > tzmax) { const float t = tzmin; tzmin = tzmax; tzmax = t; }
    if (tmin > tzmax || tzmin > tmax) { *t_near = 1e10f; *t_far = 1e10f; return; }
    if (tzmin > tmin) tmin = tzmin;
    if (tzmax < tmax) tmax = tzmax;
    *t_near = tmin > 0.0f ? tmin : 0.0f;
    *t_far = tmax;
}

__device__ __forceinline__ bool occupied(const MarchArgs& a, const float xyz[3],
                                         const uint8_t* __restrict__ grid) {
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if (xyz[d] < a.roi[d] || xyz[d] > a.roi[3 + d]) return false;
    if (!a.has_grid) return true;
    int ix[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float u = div_range(a, xyz[d] - a.roi[d], d);
        int v = (int)(u * (float)a.res);
        v = v < 0 ? 0 : (v > a.res - 1 ? a.res - 1 : v);
        ix[d] = v;
    }
    return grid[((int64_t)ix[0] * a.res + ix[1]) * a.res + ix[2]] != 0;
}

__device__ __forceinline__ float dist_next_voxel(const MarchArgs& a, const float xyz[3],
                                                 const float dir[3], const float inv[3]) {
    float t = INFINITY;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float range = a.roi[3 + d] - a.roi[d];
        const float u = div_range(a, xyz[d] - a.roi[d], d) * (float)a.res;
        const float sg = copysignf(1.0f, dir[d]);
        const float tt = div_res(a, (floorf(u + 0.5f + 0.5f * sg) - u) * inv[d]) * range;
        t = fminf(t, tt);
    }
    return fmaxf(t, 0.0f);
}

// ---- marching, G lanes to a ray ----
// The rule for one ray (nerfacc 0.3.3): from t0 = near, t1 = t0 + step,
// t_mid = (t0 + t1) / 2, while t_mid < far and fewer than cap samples (cap 0:
// no cap): the midpoint's cell occupied -> a sample [t0, t1], then t0 = t1,
// t1 = t0 + step, t_mid = (t0 + t1) / 2, and the ray ends when t_mid no
// longer grows; unoccupied -> t_mid advances by whole steps to the next voxel
// boundary (the ray ends when a step no longer advances t), t0 / t1 = t_mid
// -/+ step / 2.
// G lanes run it together.  From the group's state (t0, t1, t_mid, j), lane
// g takes g occupied steps (the same IEEE sequence) and tests its midpoint,
// so the group's grid loads are in flight together instead of one after
// another.  The first lane whose state is out of range or unoccupied ends the
// round: the lanes before it are the next samples; if it is in range, the
// voxel skip runs from its midpoint.  Samples and counts are those of the
// one-lane rule, bit for bit.  Outputs ts / te / ri may be null (count
// only); every lane returns the count.
template <int G>
__device__ int march_group(const MarchArgs& a, const float* o, const float* d, float near,
                           float far, const uint8_t* __restrict__ grid, float* __restrict__ ts,
                           float* __restrict__ te, int32_t* __restrict__ ri, int32_t ray, int g,
                           int gbase) {
    static_assert(G >= 2 && G <= 16 && (64 % G) == 0, "a group is a power of two within a wave");
    const float inv[3] = {1.0f / d[0], 1.0f / d[1], 1.0f / d[2]};
    const float step = a.step;
    const uint32_t mask = (1u << G) - 1u;
    float t0 = near;
    float t1 = t0 + step;
    float tm = (t0 + t1) * 0.5f;
    int j = 0;
    while (true) {
        // this lane's state: g occupied steps ahead
        float c0 = t0, c1 = t1, cm = tm;
        bool alive = true;
        for (int q = 0; q < g; ++q) {
            const float prev = cm;
            c0 = c1;
            c1 = c0 + step;
            cm = (c0 + c1) * 0.5f;
            alive = alive && (cm > prev);
        }
        const bool cond = alive && cm < far && !(a.cap > 0 && j + g >= a.cap);
        float xyz[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) xyz[k] = __fmaf_rn(cm, d[k], o[k]);
        const bool occ = cond && occupied(a, xyz, grid);
        const uint32_t bo = (uint32_t)(__ballot(occ) >> gbase) & mask;
        const uint32_t bc = (uint32_t)(__ballot(cond) >> gbase) & mask;
        const int m = __builtin_ctz(~bo);   // first lane not an occupied in-range state (G: none)
        if (g < m && ts) {
            ts[j + g] = c0;
            te[j + g] = c1;
            if (ri) ri[j + g] = ray;
        }
        if (m == G) {   // all G were samples: continue from the last lane's successor
            const float n0 = c1, n1 = n0 + step, nm = (n0 + n1) * 0.5f;
            const int ok = nm > cm ? 1 : 0;
            t0 = __shfl(n0, gbase + G - 1, 64);
            t1 = __shfl(n1, gbase + G - 1, 64);
            tm = __shfl(nm, gbase + G - 1, 64);
            const int okg = __shfl(ok, gbase + G - 1, 64);
            j += G;
            if (!okg) break;
            continue;
        }
        j += m;
        if (!((bc >> m) & 1u)) break;   // state m ends the march
        // state m is in range and unoccupied: the voxel skip from its midpoint
        const float tmm = __shfl(cm, gbase + m, 64);
#pragma unroll
        for (int k = 0; k < 3; ++k) xyz[k] = __fmaf_rn(tmm, d[k], o[k]);
        const float target = tmm + dist_next_voxel(a, xyz, d, inv);
        float tt = tmm;
        bool stuck = false;
        do {
            const float nt = tt + step;
            if (!(nt > tt)) {
                stuck = true;
                break;
            }
            tt = nt;
        } while (tt < target);
        if (stuck) break;
        tm = tt;
        t0 = tm - step * 0.5f;
        t1 = tm + step * 0.5f;
    }
    return j;
}

// Lanes per ray of every marching kernel (8 measured no faster).
constexpr int kMarchGroup = 4;

// ---- background marching (nerfacc 0.3.3 ray_marching with
// ContractionType.UN_BOUNDED_SPHERE and a cone angle, as called by
// models/neus.py:141-171): steps of dt = clamp(t0 * cone, step, 1e10), a
// sample kept when its midpoint's cell of the contracted grid is occupied;
// contracted scenes take no voxel skip.  [near, far] per ray come from the
// caller (its near/far planes and stratified jitter). ----
struct MarchBgArgs {
    MarchArgs m;
    float cone;
};

// Unbounded-sphere contraction of a world point into [0, 1]^3: unit cube of
// the roi -> [-1, 1]; |x| > 1 maps to (2 - 1/|x|) x/|x|; then x/4 + 1/2.
__device__ __forceinline__ bool occupied_unbounded(const MarchArgs& a, const float xyz[3],
                                                   const uint8_t* __restrict__ grid) {
    if (!a.has_grid) return true;
    float x[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[d] = div_range(a, xyz[d] - a.roi[d], d) * 2.0f - 1.0f;
    const float mag = sqrtf((x[0] * x[0] + x[1] * x[1]) + x[2] * x[2]);
    if (mag > 1.0f) {
        const float k = 2.0f - 1.0f / mag;
#pragma unroll
        for (int d = 0; d < 3; ++d) x[d] = k * (x[d] / mag);
    }
    int ix[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float u = x[d] * 0.25f + 0.5f;
        int v = (int)(u * (float)a.res);
        v = v < 0 ? 0 : (v > a.res - 1 ? a.res - 1 : v);
        ix[d] = v;
    }
    return grid[((int64_t)ix[0] * a.res + ix[1]) * a.res + ix[2]] != 0;
}

__device__ __forceinline__ float cone_dt(const MarchBgArgs& b, float t) {
    return fminf(fmaxf(t * b.cone, b.m.step), 1e10f);
}

// The background rule: from t0 = near, t1 = t0 + dt(t0), while t_mid < far
// and fewer than cap samples: the midpoint's contracted cell occupied -> a
// sample [t0, t1]; then t0 = t1, t1 = t0 + dt(t0), and the ray ends when
// t_mid no longer grows.  The steps do not depend on occupancy, so G lanes
// take G steps a round: lane g's state is g steps ahead, its sample index j
// plus the occupied lanes before it; the first state out of range (or at
// the cap) ends the ray.  Bit-exact with the one-lane rule.
template <int G>
__device__ int march_group_bg(const MarchBgArgs& b, const float* o, const float* d, float near,
                              float far, const uint8_t* __restrict__ grid, float* __restrict__ ts,
                              float* __restrict__ te, int32_t* __restrict__ ri, int32_t ray,
                              int g, int gbase) {
    const uint32_t mask = (1u << G) - 1u;
    float t0 = near;
    float t1 = t0 + cone_dt(b, t0);
    float tm = (t0 + t1) * 0.5f;
    int j = 0;
    while (true) {
        float c0 = t0, c1 = t1, cm = tm;
        bool alive = true;
        for (int q = 0; q < g; ++q) {
            const float prev = cm;
            c0 = c1;
            c1 = c0 + cone_dt(b, c0);
            cm = (c0 + c1) * 0.5f;
            alive = alive && (cm > prev);
        }
        const bool valid = alive && cm < far;
        float xyz[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) xyz[k] = __fmaf_rn(cm, d[k], o[k]);
        const bool occ = valid && occupied_unbounded(b.m, xyz, grid);
        const uint32_t bo = (uint32_t)(__ballot(occ) >> gbase) & mask;
        const int jg = j + __builtin_popcount(bo & ((1u << g) - 1u));   // this state's index
        const bool ends = !valid || (b.m.cap > 0 && jg >= b.m.cap);
        const uint32_t be = (uint32_t)(__ballot(ends) >> gbase) & mask;
        const int e = be ? __builtin_ctz(be) : G;   // first state that ends the ray
        if (g < e && occ && ts) {
            ts[jg] = c0;
            te[jg] = c1;
            ri[jg] = ray;
        }
        j += __builtin_popcount(bo & (e == G ? mask : ((1u << e) - 1u)));
        if (e < G) break;
        const float n0 = c1, n1 = n0 + cone_dt(b, n0), nm = (n0 + n1) * 0.5f;
        const int ok = nm > cm ? 1 : 0;
        t0 = __shfl(n0, gbase + G - 1, 64);
        t1 = __shfl(n1, gbase + G - 1, 64);
        tm = __shfl(nm, gbase + G - 1, 64);
        if (!__shfl(ok, gbase + G - 1, 64)) break;
    }
    return j;
}

// G lanes to a ray (whole groups share a ray and exit together).
__global__ void __launch_bounds__(64) k_march_bg_count(MarchBgArgs b,
                                                      const float* __restrict__ rays, int64_t n,
                                                      const float* __restrict__ t_near,
                                                      const float* __restrict__ t_far,
                                                      const uint8_t* __restrict__ grid,
                                                      int32_t* __restrict__ counts) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = tid / kMarchGroup;
    const int g = (int)(tid % kMarchGroup);
    if (i >= n) return;
    const float o[3] = {rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]};
    const float d[3] = {rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]};
    const int j = march_group_bg<kMarchGroup>(b, o, d, t_near[i], t_far[i], grid, nullptr, nullptr,
                                              nullptr, (int32_t)i, g, lane_id() - g);
    if (g == 0) counts[i] = j;
}

__global__ void __launch_bounds__(64) k_march_bg_fill(MarchBgArgs b, const float* __restrict__ rays,
                                                     int64_t n, const float* __restrict__ t_near,
                                                     const float* __restrict__ t_far,
                                                     const uint8_t* __restrict__ grid,
                                                     const int32_t* __restrict__ packed,
                                                     int32_t* __restrict__ ridx,
                                                     float* __restrict__ ts,
                                                     float* __restrict__ te) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = tid / kMarchGroup;
    const int g = (int)(tid % kMarchGroup);
    if (i >= n) return;
    const int32_t base = packed[2 * i];
    const float o[3] = {rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]};
    const float d[3] = {rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]};
    MarchBgArgs c = b;
    c.m.cap = packed[2 * i + 1];  // never write past the counted samples
    if (c.m.cap == 0) return;
    march_group_bg<kMarchGroup>(c, o, d, t_near[i], t_far[i], grid, ts + base, te + base,
                                ridx + base, (int32_t)i, g, lane_id() - g);
}

// Early-stop visibility of the background samples (nerfacc 0.3.3
// ray_marching with a sigma_fn, alpha_thre 0, as models/neus.py:157-167
// calls it): alpha = 1 - exp(-sigma (t1 - t0)), T_j = prod_{i<j} (1 - alpha_i)
// in sample order, a sample kept while T_j >= eps.  T never grows along a
// ray, so the kept samples are a prefix of the ray's; one lane per ray walks
// it sequentially (the product's rounding order is the sample order).
__global__ void k_visible_prefix(const float* __restrict__ sigma, const float* __restrict__ ts,
                                 const float* __restrict__ te, const int32_t* __restrict__ packed,
                                 int64_t n_rays, float eps, int32_t* __restrict__ keep) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rays) return;
    const int32_t b = packed[2 * r], c = packed[2 * r + 1];
    float T = 1.0f;
    int32_t k = 0;
    for (; k < c; ++k) {
        if (!(T >= eps)) break;
        const float alpha = 1.0f - expf(-(sigma[b + k] * (te[b + k] - ts[b + k])));
        T = T * (1.0f - alpha);
    }
    keep[r] = k;
}

__device__ __forceinline__ float ray_near(const float* t_near, const float* jitter, int64_t i,
                                          float step) {
    float nr = t_near[i];
    if (jitter) nr = nr + jitter[i] * step;
    return nr;
}

__global__ void k_ray_aabb(const float* __restrict__ rays, int64_t n, float aabb0, float aabb1,
                           float aabb2, float aabb3, float aabb4, float aabb5,
                           float* __restrict__ t_near, float* __restrict__ t_far) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float aabb[6] = {aabb0, aabb1, aabb2, aabb3, aabb4, aabb5};
    float o[3] = {rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]};
    float d[3] = {rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]};
    aabb_one(o, d, aabb, t_near + i, t_far + i);
}

__global__ void __launch_bounds__(64) k_march_count(MarchArgs a, const float* __restrict__ rays,
                                                   int64_t n, const float* __restrict__ t_near,
                                                   const float* __restrict__ t_far,
                                                   const float* __restrict__ jitter,
                                                   const uint8_t* __restrict__ grid,
                                                   int32_t* __restrict__ counts) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = tid / kMarchGroup;
    const int g = (int)(tid % kMarchGroup);
    if (i >= n) return;
    const float o[3] = {rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]};
    const float d[3] = {rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]};
    const int j = march_group<kMarchGroup>(a, o, d, ray_near(t_near, jitter, i, a.step), t_far[i],
                                           grid, nullptr, nullptr, nullptr, (int32_t)i, g,
                                           lane_id() - g);
    if (g == 0) counts[i] = j;
}

// packed[i] = (start, count) from counts and their exclusive scan; total.
__global__ void k_march_pack(const int32_t* __restrict__ counts, const int32_t* __restrict__ starts,
                             int64_t n, int32_t* __restrict__ packed, int32_t* __restrict__ total) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n == 0 && i == 0) total[0] = 0;
    if (i >= n) return;
    packed[2 * i] = starts[i];
    packed[2 * i + 1] = counts[i];
    if (i == n - 1) total[0] = starts[i] + counts[i];
}

__global__ void __launch_bounds__(64) k_march_fill(MarchArgs a, const float* __restrict__ rays,
                                                  int64_t n, const float* __restrict__ t_near,
                                                  const float* __restrict__ t_far,
                                                  const float* __restrict__ jitter,
                                                  const uint8_t* __restrict__ grid,
                                                  const int32_t* __restrict__ packed,
                                                  int32_t* __restrict__ ridx,
                                                  float* __restrict__ ts, float* __restrict__ te) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = tid / kMarchGroup;
    const int g = (int)(tid % kMarchGroup);
    if (i >= n) return;
    const int32_t base = packed[2 * i];
    const float o[3] = {rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]};
    const float d[3] = {rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]};
    MarchArgs b = a;
    b.cap = packed[2 * i + 1];  // never write past the counted samples
    if (b.cap == 0) return;
    march_group<kMarchGroup>(b, o, d, ray_near(t_near, jitter, i, a.step), t_far[i], grid,
                             ts + base, te + base, ridx + base, (int32_t)i, g, lane_id() - g);
}

// Staged marching (a per-ray sample cap is set): the count pass also writes
// each ray's samples into a cap-strided staging area, and the fill pass
// copies them to their packed positions instead of marching again.
__global__ void __launch_bounds__(64) k_march_count_staged(MarchArgs a,
                                                          const float* __restrict__ rays, int64_t n,
                                                          const float* __restrict__ t_near,
                                                          const float* __restrict__ t_far,
                                                          const float* __restrict__ jitter,
                                                          const uint8_t* __restrict__ grid,
                                                          int32_t* __restrict__ counts,
                                                          float* __restrict__ st0,
                                                          float* __restrict__ st1) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = tid / kMarchGroup;
    const int g = (int)(tid % kMarchGroup);
    if (i >= n) return;
    const float o[3] = {rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]};
    const float d[3] = {rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]};
    const int64_t base = i * (int64_t)a.cap;   // a.cap > 0: samples j < cap
    const int j = march_group<kMarchGroup>(a, o, d, ray_near(t_near, jitter, i, a.step), t_far[i],
                                           grid, st0 + base, st1 + base, nullptr, (int32_t)i, g,
                                           lane_id() - g);
    if (g == 0) counts[i] = j;
}

// One wave per ray: its staged samples to the packed outputs, coalesced.
__global__ void __launch_bounds__(256) k_march_fill_staged(const int32_t* __restrict__ packed,
                                                           int64_t n, int32_t cap,
                                                           const float* __restrict__ st0,
                                                           const float* __restrict__ st1,
                                                           int32_t* __restrict__ ridx,
                                                           float* __restrict__ ts,
                                                           float* __restrict__ te) {
    const int64_t r = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (r >= n) return;
    const int l = threadIdx.x & 63;
    const int32_t base = packed[2 * r], cnt = packed[2 * r + 1];
    const int64_t src = r * (int64_t)cap;
    for (int j = l; j < cnt; j += 64) {
        ts[base + j] = st0[src + j];
        te[base + j] = st1[src + j];
        ridx[base + j] = (int32_t)r;
    }
}

static MarchArgs march_args(const float* roi6, const uint8_t* grid, int res, float step, int cap) {
    MarchArgs a;
    for (int d = 0; d < 6; ++d) a.roi[d] = roi6[d];
    a.res = res;
    a.step = step;
    a.cap = cap;
    a.has_grid = grid != nullptr;
    a.range_pow2 = 0;
    for (int d = 0; d < 3; ++d) {
        const float range = roi6[3 + d] - roi6[d];
        int ex = 0;
        const float m = std::frexp(range, &ex);   // range = m * 2^ex, m in [0.5, 1)
        const bool pow2 = std::isfinite(range) && range > 0.0f && m == 0.5f &&
                          std::isnormal(1.0f / range);
        a.inv_range[d] = pow2 ? 1.0f / range : 0.0f;
        if (pow2) a.range_pow2 |= 1 << d;
    }
    a.res_pow2 = res > 0 && (res & (res - 1)) == 0;
    a.inv_res = a.res_pow2 ? 1.0f / (float)res : 0.0f;
    return a;
}

static size_t scan_tmp_bytes(int64_t n) {
    size_t bytes = 0;
    rocprim::exclusive_scan(nullptr, bytes, (const int32_t*)nullptr, (int32_t*)nullptr, 0,
                            (size_t)n, rocprim::plus<int32_t>(), 0, false);
    return bytes;
}

}  // namespace am

using namespace am;

extern "C" int am_ray_aabb(const float* rays, int64_t n_rays, const float* aabb6_host,
                           float* t_near, float* t_far, void* stream) {
    AM_CHECK_ARG(n_rays >= 0, "n_rays < 0");
    if (n_rays == 0) return AM_OK;
    AM_CHECK_ARG(rays && aabb6_host && t_near && t_far, "NULL buffer");
    const float* b = aabb6_host;
    hipLaunchKernelGGL(k_ray_aabb, dim3((unsigned)ceil_div(n_rays, 256)), dim3(256), 0,
                       to_stream(stream), rays, n_rays, b[0], b[1], b[2], b[3], b[4], b[5], t_near,
                       t_far);
    AM_LAUNCH_CHECK();
    return AM_OK;
}

extern "C" size_t am_march_workspace_size(int64_t n_rays) {
    if (n_rays < 0) return 0;
    WsCarver c(nullptr, 0);
    c.take<int32_t>((size_t)n_rays + 1);
    c.take<int32_t>((size_t)n_rays + 1);
    c.take<char>(scan_tmp_bytes(n_rays > 0 ? n_rays : 1) + 256);
    return c.used + 256;
}

static int check_march(int64_t n_rays, const float* rays, const float* t_near, const float* t_far,
                       const float* roi6, const uint8_t* grid, int32_t res, float step,
                       int32_t cap) {
    AM_CHECK_ARG(n_rays >= 0 && n_rays < (1ll << 31), "n_rays %lld", (long long)n_rays);
    AM_CHECK_ARG(rays && t_near && t_far && roi6, "NULL buffer");
    AM_CHECK_ARG(!grid || (res >= 1 && res <= 4096), "grid_res %d", res);
    AM_CHECK_ARG(step > 0.0f, "step %g <= 0", (double)step);
    AM_CHECK_ARG(cap >= 0, "cap < 0");
    for (int d = 0; d < 3; ++d) AM_CHECK_ARG(roi6[3 + d] > roi6[d], "empty roi");
    return AM_OK;
}

extern "C" int am_march_count(const float* rays, int64_t n_rays, const float* t_near,
                              const float* t_far, const float* jitter, const float* roi6_host,
                              const uint8_t* grid, int32_t grid_res, float step, int32_t cap,
                              int32_t* packed, int32_t* total, void* workspace,
                              size_t workspace_bytes, void* stream) {
    if (int rc = check_march(n_rays, rays, t_near, t_far, roi6_host, grid, grid_res, step, cap))
        return rc;
    AM_CHECK_ARG(packed && total && workspace, "NULL packed/total/workspace");
    hipStream_t s = to_stream(stream);
    if (n_rays == 0) {
        hipLaunchKernelGGL(k_march_pack, dim3(1), dim3(1), 0, s, (const int32_t*)nullptr,
                           (const int32_t*)nullptr, (int64_t)0, packed, total);
        AM_LAUNCH_CHECK();
        return AM_OK;
    }
    WsCarver c(workspace, workspace_bytes);
    int32_t* counts = c.take<int32_t>((size_t)n_rays + 1);
    int32_t* starts = c.take<int32_t>((size_t)n_rays + 1);
    size_t sb = scan_tmp_bytes(n_rays);
    char* stmp = c.take<char>(sb + 256);
    if (!c.ok()) return workspace_short(c.used, workspace_bytes);
    const MarchArgs a = march_args(roi6_host, grid, grid_res, step, cap);
    const unsigned nb = (unsigned)ceil_div(n_rays * kMarchGroup, 64);
    {
        KernelTimer kt_("march_count", s);
        hipLaunchKernelGGL(k_march_count, dim3(nb), dim3(64), 0, s, a, rays, n_rays, t_near, t_far,
                           jitter, grid, counts);
    }
    AM_LAUNCH_CHECK();
    hipError_t e = rocprim::exclusive_scan(stmp, sb, counts, starts, 0, (size_t)n_rays,
                                           rocprim::plus<int32_t>(), s, false);
    if (e != hipSuccess) {
        set_error("exclusive scan failed: %s", hipGetErrorString(e));
        return AM_ERR_HIP;
    }
    hipLaunchKernelGGL(k_march_pack, dim3((unsigned)ceil_div(n_rays, 256)), dim3(256), 0, s, counts,
                       starts, n_rays, packed, total);
    AM_LAUNCH_CHECK();
    return AM_OK;
}

extern "C" int am_march_fill(const float* rays, int64_t n_rays, const float* t_near,
                             const float* t_far, const float* jitter, const float* roi6_host,
                             const uint8_t* grid, int32_t grid_res, float step, int32_t cap,
                             const int32_t* packed, int32_t* ray_indices, float* t_starts,
                             float* t_ends, void* stream) {
    if (int rc = check_march(n_rays, rays, t_near, t_far, roi6_host, grid, grid_res, step, cap))
        return rc;
    if (n_rays == 0) return AM_OK;
    AM_CHECK_ARG(packed && ray_indices && t_starts && t_ends, "NULL output");
    const MarchArgs a = march_args(roi6_host, grid, grid_res, step, cap);
    {
        KernelTimer kt_("march_fill", to_stream(stream));
        hipLaunchKernelGGL(k_march_fill, dim3((unsigned)ceil_div(n_rays * kMarchGroup, 64)), dim3(64), 0,
                           to_stream(stream), a, rays, n_rays, t_near, t_far, jitter, grid, packed,
                           ray_indices, t_starts, t_ends);
    }
    AM_LAUNCH_CHECK();
    return AM_OK;
}

extern "C" size_t am_march_staged_workspace_size(int64_t n_rays, int32_t cap) {
    if (n_rays < 0 || cap <= 0) return 0;
    WsCarver c(nullptr, 0);
    c.take<int32_t>((size_t)n_rays + 1);
    c.take<int32_t>((size_t)n_rays + 1);
    c.take<char>(scan_tmp_bytes(n_rays > 0 ? n_rays : 1) + 256);
    c.take<float>((size_t)n_rays * (size_t)cap);
    c.take<float>((size_t)n_rays * (size_t)cap);
    return c.used + 256;
}

extern "C" int am_march_count_staged(const float* rays, int64_t n_rays, const float* t_near,
                                     const float* t_far, const float* jitter,
                                     const float* roi6_host, const uint8_t* grid, int32_t grid_res,
                                     float step, int32_t cap, int32_t* packed, int32_t* total,
                                     void* workspace, size_t workspace_bytes, void* stream) {
    if (int rc = check_march(n_rays, rays, t_near, t_far, roi6_host, grid, grid_res, step, cap))
        return rc;
    AM_CHECK_ARG(cap > 0, "staged marching needs a per-ray sample cap (got %d)", cap);
    AM_CHECK_ARG(packed && total && workspace, "NULL packed/total/workspace");
    hipStream_t s = to_stream(stream);
    if (n_rays == 0) {
        hipLaunchKernelGGL(k_march_pack, dim3(1), dim3(1), 0, s, (const int32_t*)nullptr,
                           (const int32_t*)nullptr, (int64_t)0, packed, total);
        AM_LAUNCH_CHECK();
        return AM_OK;
    }
    WsCarver c(workspace, workspace_bytes);
    int32_t* counts = c.take<int32_t>((size_t)n_rays + 1);
    int32_t* starts = c.take<int32_t>((size_t)n_rays + 1);
    size_t sb = scan_tmp_bytes(n_rays);
    char* stmp = c.take<char>(sb + 256);
    float* st0 = c.take<float>((size_t)n_rays * (size_t)cap);
    float* st1 = c.take<float>((size_t)n_rays * (size_t)cap);
    if (!c.ok()) return workspace_short(c.used, workspace_bytes);
    const MarchArgs a = march_args(roi6_host, grid, grid_res, step, cap);
    {
        KernelTimer kt_("march_count", s);
        hipLaunchKernelGGL(k_march_count_staged,
                           dim3((unsigned)ceil_div(n_rays * kMarchGroup, 64)), dim3(64), 0, s, a,
                           rays, n_rays, t_near, t_far, jitter, grid, counts, st0, st1);
    }
    AM_LAUNCH_CHECK();
    hipError_t e = rocprim::exclusive_scan(stmp, sb, counts, starts, 0, (size_t)n_rays,
                                           rocprim::plus<int32_t>(), s, false);
    if (e != hipSuccess) {
        set_error("exclusive scan failed: %s", hipGetErrorString(e));
        return AM_ERR_HIP;
    }
    hipLaunchKernelGGL(k_march_pack, dim3((unsigned)ceil_div(n_rays, 256)), dim3(256), 0, s, counts,
                       starts, n_rays, packed, total);
    AM_LAUNCH_CHECK();
    return AM_OK;
}

extern "C" int am_march_fill_staged(int64_t n_rays, int32_t cap, const int32_t* packed,
                                    const void* workspace, size_t workspace_bytes,
                                    int32_t* ray_indices, float* t_starts, float* t_ends,
                                    void* stream) {
    AM_CHECK_ARG(n_rays >= 0 && n_rays < (1ll << 31), "n_rays %lld", (long long)n_rays);
    AM_CHECK_ARG(cap > 0, "staged marching needs a per-ray sample cap (got %d)", cap);
    if (n_rays == 0) return AM_OK;
    AM_CHECK_ARG(packed && workspace && ray_indices && t_starts && t_ends, "NULL buffer");
    // the same carve as am_march_count_staged
    WsCarver c(const_cast<void*>(workspace), workspace_bytes);
    c.take<int32_t>((size_t)n_rays + 1);
    c.take<int32_t>((size_t)n_rays + 1);
    c.take<char>(scan_tmp_bytes(n_rays) + 256);
    const float* st0 = c.take<float>((size_t)n_rays * (size_t)cap);
    const float* st1 = c.take<float>((size_t)n_rays * (size_t)cap);
    if (!c.ok()) return workspace_short(c.used, workspace_bytes);
    {
        KernelTimer kt_("march_fill", to_stream(stream));
        hipLaunchKernelGGL(k_march_fill_staged, dim3((unsigned)ceil_div(n_rays, 4)), dim3(256), 0,
                           to_stream(stream), packed, n_rays, cap, st0, st1, ray_indices, t_starts,
                           t_ends);
    }
    AM_LAUNCH_CHECK();
    return AM_OK;
}

extern "C" int am_march_bg_count(const float* rays, int64_t n_rays, const float* t_near,
                                 const float* t_far, const float* roi6_host, const uint8_t* grid,
                                 int32_t grid_res, float step, float cone_angle, int32_t cap,
                                 int32_t* packed, int32_t* total, void* workspace,
                                 size_t workspace_bytes, void* stream) {
    if (int rc = check_march(n_rays, rays, t_near, t_far, roi6_host, grid, grid_res, step, cap))
        return rc;
    AM_CHECK_ARG(cone_angle >= 0.0f, "cone_angle %g < 0", (double)cone_angle);
    AM_CHECK_ARG(packed && total && workspace, "NULL output");
    hipStream_t s = to_stream(stream);
    if (n_rays == 0) {
        hipLaunchKernelGGL(k_march_pack, dim3(1), dim3(1), 0, s, (const int32_t*)nullptr,
                           (const int32_t*)nullptr, (int64_t)0, packed, total);
        AM_LAUNCH_CHECK();
        return AM_OK;
    }
    WsCarver c(workspace, workspace_bytes);
    int32_t* counts = c.take<int32_t>((size_t)n_rays + 1);
    int32_t* starts = c.take<int32_t>((size_t)n_rays + 1);
    size_t sb = scan_tmp_bytes(n_rays);
    char* stmp = c.take<char>(sb + 256);
    if (!c.ok()) return workspace_short(c.used, workspace_bytes);
    MarchBgArgs b;
    b.m = march_args(roi6_host, grid, grid_res, step, cap);
    b.cone = cone_angle;
    {
        KernelTimer kt_("march_bg_count", s);
        hipLaunchKernelGGL(k_march_bg_count, dim3((unsigned)ceil_div(n_rays * kMarchGroup, 64)),
                           dim3(64), 0, s,
                           b, rays, n_rays, t_near, t_far, grid, counts);
    }
    AM_LAUNCH_CHECK();
    hipError_t e = rocprim::exclusive_scan(stmp, sb, counts, starts, 0, (size_t)n_rays,
                                           rocprim::plus<int32_t>(), s, false);
    if (e != hipSuccess) {
        set_error("exclusive scan failed: %s", hipGetErrorString(e));
        return AM_ERR_HIP;
    }
    hipLaunchKernelGGL(k_march_pack, dim3((unsigned)ceil_div(n_rays, 256)), dim3(256), 0, s, counts,
                       starts, n_rays, packed, total);
    AM_LAUNCH_CHECK();
    return AM_OK;
}

extern "C" int am_march_bg_fill(const float* rays, int64_t n_rays, const float* t_near,
                                const float* t_far, const float* roi6_host, const uint8_t* grid,
                                int32_t grid_res, float step, float cone_angle, int32_t cap,
                                const int32_t* packed, int32_t* ray_indices, float* t_starts,
                                float* t_ends, void* stream) {
    if (int rc = check_march(n_rays, rays, t_near, t_far, roi6_host, grid, grid_res, step, cap))
        return rc;
    AM_CHECK_ARG(cone_angle >= 0.0f, "cone_angle %g < 0", (double)cone_angle);
    if (n_rays == 0) return AM_OK;
    AM_CHECK_ARG(packed && ray_indices && t_starts && t_ends, "NULL output");
    MarchBgArgs b;
    b.m = march_args(roi6_host, grid, grid_res, step, cap);
    b.cone = cone_angle;
    {
        KernelTimer kt_("march_bg_fill", to_stream(stream));
        hipLaunchKernelGGL(k_march_bg_fill, dim3((unsigned)ceil_div(n_rays * kMarchGroup, 64)),
                           dim3(64), 0,
                           to_stream(stream), b, rays, n_rays, t_near, t_far, grid, packed,
                           ray_indices, t_starts, t_ends);
    }
    AM_LAUNCH_CHECK();
    return AM_OK;
}

extern "C" int am_march_bg_visible(const float* sigma, const float* t_starts, const float* t_ends,
                                   const int32_t* packed, int64_t n_rays, float early_stop_eps,
                                   int32_t* packed_out, int32_t* total, void* workspace,
                                   size_t workspace_bytes, void* stream) {
    AM_CHECK_ARG(n_rays >= 0 && n_rays < (1ll << 31), "n_rays %lld", (long long)n_rays);
    AM_CHECK_ARG(packed_out && total, "NULL packed_out/total");
    AM_CHECK_ARG(early_stop_eps >= 0.0f, "early_stop_eps %g < 0", (double)early_stop_eps);
    hipStream_t s = to_stream(stream);
    if (n_rays == 0) {
        hipLaunchKernelGGL(k_march_pack, dim3(1), dim3(1), 0, s, (const int32_t*)nullptr,
                           (const int32_t*)nullptr, (int64_t)0, packed_out, total);
        AM_LAUNCH_CHECK();
        return AM_OK;
    }
    AM_CHECK_ARG(sigma && t_starts && t_ends && packed && workspace, "NULL buffer");
    WsCarver c(workspace, workspace_bytes);
    int32_t* keep = c.take<int32_t>((size_t)n_rays + 1);
    int32_t* starts = c.take<int32_t>((size_t)n_rays + 1);
    size_t sb = scan_tmp_bytes(n_rays);
    char* stmp = c.take<char>(sb + 256);
    if (!c.ok()) return workspace_short(c.used, workspace_bytes);
    {
        KernelTimer kt_("march_bg_visible", s);
        hipLaunchKernelGGL(k_visible_prefix, dim3((unsigned)ceil_div(n_rays, 64)), dim3(64), 0, s,
                           sigma, t_starts, t_ends, packed, n_rays, early_stop_eps, keep);
    }
    AM_LAUNCH_CHECK();
    hipError_t e = rocprim::exclusive_scan(stmp, sb, keep, starts, 0, (size_t)n_rays,
                                           rocprim::plus<int32_t>(), s, false);
    if (e != hipSuccess) {
        set_error("exclusive scan failed: %s", hipGetErrorString(e));
        return AM_ERR_HIP;
    }
    hipLaunchKernelGGL(k_march_pack, dim3((unsigned)ceil_div(n_rays, 256)), dim3(256), 0, s, keep,
                       starts, n_rays, packed_out, total);
    AM_LAUNCH_CHECK();
    return AM_OK;
}
