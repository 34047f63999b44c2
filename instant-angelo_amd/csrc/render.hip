// Sample positions and NeuS compositing: alpha (models/neus.py:117-139),
// transmittance / weights (nerfacc render_weight_from_alpha, :234) and the
// per-ray sums of accumulate_along_rays (:235-241), one wavefront per ray.
//
// Transmittance T_i = prod_{j<i} (1 - a_j) is a segmented multiplicative scan
// over the ray's samples in chunks of 64 lanes with a running carry.  Its
// backward uses dL/da_i = T_i * (g_i - R_i), R_i = sum_{j>i} g_j a_j
// prod_{i<k<j} (1 - a_k), evaluated as a reverse scan of the affine maps
// R <- (1-a_j) R + g_j a_j, so no division by (1 - a_i) is needed.
#include "am_common.h"
#include "table_grad.h"
#include <mutex>

namespace am {

// ---- alpha (torch op order of NeuSModel.get_alpha) ----
struct AlphaCtx {
    float t1, ic, prev, next, pcdf, ncdf, num, den, raw, alpha;
};

__device__ __forceinline__ float sigmoidf_(float z) { return 1.0f / (1.0f + expf(-z)); }

__device__ __forceinline__ AlphaCtx neus_alpha(float sdf, float cosv, float dist, float inv_s,
                                               float one_m_ratio, float ratio) {
    AlphaCtx a;
    a.t1 = (-cosv) * 0.5f + 0.5f;
    const float r1 = a.t1 > 0.0f ? a.t1 : 0.0f;
    const float r2 = -cosv > 0.0f ? -cosv : 0.0f;
    a.ic = -(r1 * one_m_ratio + r2 * ratio);
    const float h = (a.ic * dist) * 0.5f;
    a.next = sdf + h;
    a.prev = sdf - h;
    a.pcdf = sigmoidf_(a.prev * inv_s);
    a.ncdf = sigmoidf_(a.next * inv_s);
    a.num = (a.pcdf - a.ncdf) + 1e-5f;
    a.den = a.pcdf + 1e-5f;
    a.raw = a.num / a.den;
    a.alpha = fminf(fmaxf(a.raw, 0.0f), 1.0f);
    return a;
}

// Gradients of alpha w.r.t. sdf, cos and inv_s.
__device__ __forceinline__ void neus_alpha_bwd(const AlphaCtx& a, float g_alpha, float cosv,
                                               float dist, float inv_s, float one_m_ratio,
                                               float ratio, float* g_sdf, float* g_cos,
                                               float* g_inv_s) {
    const float g_raw = (a.raw >= 0.0f && a.raw <= 1.0f) ? g_alpha : 0.0f;
    const float g_num = g_raw / a.den;
    const float g_den = -g_raw * a.num / (a.den * a.den);
    const float g_pc = g_num + g_den;
    const float g_nc = -g_num;
    const float g_zp = g_pc * a.pcdf * (1.0f - a.pcdf);
    const float g_zn = g_nc * a.ncdf * (1.0f - a.ncdf);
    const float g_prev = g_zp * inv_s, g_next = g_zn * inv_s;
    *g_inv_s = g_zp * a.prev + g_zn * a.next;
    *g_sdf = g_prev + g_next;
    const float g_ic = (g_next - g_prev) * dist * 0.5f;
    const float dic = (a.t1 > 0.0f ? 0.5f * one_m_ratio : 0.0f) + (-cosv > 0.0f ? ratio : 0.0f);
    *g_cos = g_ic * dic;
}

// ---- wave primitives ----
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Inclusive product scan over lanes.
__device__ __forceinline__ float wave_prod_scan(float v) {
    const int l = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const float u = __shfl_up(v, o, 64);
        if (l >= o) v *= u;
    }
    return v;
}

// Compensated product (Dekker/FMA two-product): returns hi, accumulates lo.
__device__ __forceinline__ float two_prod(float a, float b, float* err) {
    const float p = a * b;
    *err = __fmaf_rn(a, b, -p);
    return p;
}

// Kahan-Babuska (Neumaier) accumulation.
struct KSum {
    float s = 0.0f, c = 0.0f;
    __device__ __forceinline__ void add(float x) {
        const float t = s + x;
        if (fabsf(s) >= fabsf(x)) c += (s - t) + x;
        else c += (x - t) + s;
        s = t;
    }
    __device__ __forceinline__ float value() const { return s + c; }
};

__device__ __forceinline__ void load3(const float* p, int64_t i, float v[3]) {
    v[0] = p[3 * i];
    v[1] = p[3 * i + 1];
    v[2] = p[3 * i + 2];
}

__global__ void k_sample_points(const float* __restrict__ rays, const int32_t* __restrict__ ridx,
                                const float* __restrict__ t0, const float* __restrict__ t1,
                                int64_t n, float* __restrict__ pts, float* __restrict__ mid,
                                float* __restrict__ dists) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float a = t0[i], b = t1[i];
    const float m = (a + b) * 0.5f;
    const int64_t r = ridx[i];
    if (pts) {
#pragma unroll
        for (int d = 0; d < 3; ++d) pts[3 * i + d] = rays[6 * r + d] + rays[6 * r + 3 + d] * m;
    }
    if (mid) mid[i] = m;
    if (dists) dists[i] = b - a;
}

struct CompArgs {
    float ratio, one_m_ratio;
    int kahan;
};

// One wave per ray.
__global__ void __launch_bounds__(256) k_composite_fwd(
    CompArgs ca, const float* __restrict__ sdf, const float* __restrict__ normal,
    const float* __restrict__ rays, const float* __restrict__ t0, const float* __restrict__ t1,
    const float* __restrict__ rgb, const float* __restrict__ inv_s_p,
    const int32_t* __restrict__ packed, int64_t n_rays, float* __restrict__ weights,
    float* __restrict__ trans, float* __restrict__ opacity, float* __restrict__ depth,
    float* __restrict__ comp_rgb, float* __restrict__ comp_normal,
    float* __restrict__ comp_normal_raw) {
    const int64_t r = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (r >= n_rays) return;
    const int l = lane_id();
    const int start = packed[2 * r], count = packed[2 * r + 1];
    const float inv_s = inv_s_p[0];
    float dir[3];
    load3(rays + 3, 2 * r, dir);  // rays[6r+3 .. 6r+5]
    float T = 1.0f, Tlo = 0.0f;
    KSum s_op, s_dep, s_rgb[3], s_n[3];
    for (int c0 = 0; c0 < count; c0 += 64) {
        const bool valid = c0 + l < count;
        const int64_t j = (int64_t)start + c0 + l;
        float alpha = 0.0f, w = 0.0f, m = 0.0f, col[3] = {0, 0, 0}, nn[3] = {0, 0, 0};
        if (valid) {
            load3(normal, j, nn);
            const float cosv = (dir[0] * nn[0] + dir[1] * nn[1]) + dir[2] * nn[2];
            const float a = t0[j], b = t1[j];
            m = (a + b) * 0.5f;
            const AlphaCtx ac = neus_alpha(sdf[j], cosv, b - a, inv_s, ca.one_m_ratio, ca.ratio);
            alpha = ac.alpha;
            load3(rgb, j, col);
        }
        const float om = 1.0f - alpha;
        const float incl = wave_prod_scan(om);
        float excl = __shfl_up(incl, 1, 64);
        if (l == 0) excl = 1.0f;
        float Ti;
        if (ca.kahan) {
            float e;
            Ti = two_prod(T, excl, &e);
            Ti = Ti + (e + Tlo * excl);
        } else {
            Ti = T * excl;
        }
        w = Ti * alpha;
        if (valid) {
            weights[j] = w;
            if (trans) trans[j] = Ti;
        }
        const float v_op = w, v_dep = w * m;
        const float v_rgb[3] = {w * col[0], w * col[1], w * col[2]};
        const float v_n[3] = {w * nn[0], w * nn[1], w * nn[2]};
        s_op.add(wave_sum(v_op));
        s_dep.add(wave_sum(v_dep));
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            s_rgb[d].add(wave_sum(v_rgb[d]));
            s_n[d].add(wave_sum(v_n[d]));
        }
        const float chunk = __shfl(incl, 63, 64);
        if (ca.kahan) {
            float e;
            const float hi = two_prod(T, chunk, &e);
            Tlo = e + Tlo * chunk;
            T = hi;
        } else {
            T = T * chunk;
        }
    }
    if (l == 0) {
        const bool kh = ca.kahan != 0;
        const float op = kh ? s_op.value() : s_op.s;
        opacity[r] = op;
        depth[r] = kh ? s_dep.value() : s_dep.s;
        float an[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            comp_rgb[3 * r + d] = kh ? s_rgb[d].value() : s_rgb[d].s;
            an[d] = kh ? s_n[d].value() : s_n[d].s;
            if (comp_normal_raw) comp_normal_raw[3 * r + d] = an[d];
        }
        const float nrm = sqrtf((an[0] * an[0] + an[1] * an[1]) + an[2] * an[2]);
        const float dnm = fmaxf(nrm, 1e-12f);
        const bool fg = op > 0.1f;
#pragma unroll
        for (int d = 0; d < 3; ++d) comp_normal[3 * r + d] = fg ? an[d] / dnm : 0.0f;
    }
}

// Composes affine maps f(R) = A*R + B (outer o inner).
__device__ __forceinline__ void affine_compose(float Ao, float Bo, float Ai, float Bi, float* A,
                                               float* B) {
    *A = Ao * Ai;
    *B = Ao * Bi + Bo;
}

__global__ void __launch_bounds__(256) k_composite_bwd(
    CompArgs ca, const float* __restrict__ sdf, const float* __restrict__ normal,
    const float* __restrict__ rays, const float* __restrict__ t0, const float* __restrict__ t1,
    const float* __restrict__ rgb, const float* __restrict__ inv_s_p,
    const int32_t* __restrict__ packed, int64_t n_rays, const float* __restrict__ trans,
    const float* __restrict__ opacity, const float* __restrict__ cn_raw,
    const float* __restrict__ g_weights, const float* __restrict__ g_opacity,
    const float* __restrict__ g_depth, const float* __restrict__ g_comp_rgb,
    const float* __restrict__ g_comp_normal, float* __restrict__ g_sdf,
    float* __restrict__ g_normal, float* __restrict__ g_rgb, float* __restrict__ g_inv_s_ray) {
    const int64_t r = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (r >= n_rays) return;
    const int l = lane_id();
    const int start = packed[2 * r], count = packed[2 * r + 1];
    const float inv_s = inv_s_p[0];
    float dir[3];
    load3(rays + 3, 2 * r, dir);
    // per-ray upstream gradients
    const float gop = g_opacity ? g_opacity[r] : 0.0f;
    const float gdep = g_depth ? g_depth[r] : 0.0f;
    float gcol[3] = {0, 0, 0}, gacc_n[3] = {0, 0, 0};
    if (g_comp_rgb) load3(g_comp_rgb, r, gcol);
    if (g_comp_normal && opacity[r] > 0.1f) {
        // comp_normal = an / max(|an|, 1e-12)
        float an[3], gn[3];
        load3(cn_raw, r, an);
        load3(g_comp_normal, r, gn);
        const float nrm = sqrtf((an[0] * an[0] + an[1] * an[1]) + an[2] * an[2]);
        const float dnm = fmaxf(nrm, 1e-12f);
        const float dot = gn[0] * an[0] + gn[1] * an[1] + gn[2] * an[2];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            float v = gn[d] / dnm;
            if (nrm >= 1e-12f && nrm > 0.0f) v -= dot / (dnm * dnm) * (an[d] / nrm);
            gacc_n[d] = v;
        }
    }
    float Rcarry = 0.0f;
    float ginv_acc = 0.0f;
    const int nchunks = (count + 63) / 64;
    for (int ch = nchunks - 1; ch >= 0; --ch) {
        const int c0 = ch * 64;
        const bool valid = c0 + l < count;
        const int64_t j = (int64_t)start + c0 + l;
        float nn[3] = {0, 0, 0}, col[3] = {0, 0, 0}, m = 0.0f, dist = 0.0f, cosv = 0.0f;
        AlphaCtx ac{};
        float gw = 0.0f, Ti = 0.0f;
        if (valid) {
            load3(normal, j, nn);
            load3(rgb, j, col);
            cosv = (dir[0] * nn[0] + dir[1] * nn[1]) + dir[2] * nn[2];
            const float a = t0[j], b = t1[j];
            m = (a + b) * 0.5f;
            dist = b - a;
            ac = neus_alpha(sdf[j], cosv, dist, inv_s, ca.one_m_ratio, ca.ratio);
            Ti = trans[j];
            gw = gop + gdep * m + (gcol[0] * col[0] + gcol[1] * col[1] + gcol[2] * col[2]) +
                 (gacc_n[0] * nn[0] + gacc_n[1] * nn[1] + gacc_n[2] * nn[2]);
            if (g_weights) gw += g_weights[j];
        }
        const float alpha = valid ? ac.alpha : 0.0f;
        // inclusive reverse scan of f_j(R) = (1-a_j) R + g_j a_j
        float A = 1.0f - alpha, B = gw * alpha;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const float Ad = __shfl_down(A, o, 64), Bd = __shfl_down(B, o, 64);
            if (l + o < 64) affine_compose(A, B, Ad, Bd, &A, &B);
        }
        // exclusive (maps after j within the chunk), applied to the carry
        float An = __shfl_down(A, 1, 64), Bn = __shfl_down(B, 1, 64);
        if (l == 63) { An = 1.0f; Bn = 0.0f; }
        const float Ri = An * Rcarry + Bn;
        const float A0 = __shfl(A, 0, 64), B0 = __shfl(B, 0, 64);
        Rcarry = A0 * Rcarry + B0;
        if (valid) {
            const float g_alpha = Ti * (gw - Ri);
            float gs, gc, gi;
            neus_alpha_bwd(ac, g_alpha, cosv, dist, inv_s, ca.one_m_ratio, ca.ratio, &gs, &gc, &gi);
            const float w = Ti * alpha;
            g_sdf[j] = gs;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                g_normal[3 * j + d] = gc * dir[d] + w * gacc_n[d];
                g_rgb[3 * j + d] = w * gcol[d];
            }
            ginv_acc += gi;
        }
    }
    const float tot = wave_sum(ginv_acc);
    if (l == 0) g_inv_s_ray[r] = tot;
}

// ---- background compositing: nerfacc render_weight_from_density +
// accumulate_along_rays as called by models/neus.py:174-178.  s_i =
// sigma_i (t1_i - t0_i), T_i = exp(-sum_{j<i} s_j), w_i = T_i (1 - exp(-s_i));
// opacity = sum w, depth = sum w mid, rgb = sum w c.  One wave per ray, the
// exclusive sum of s as a segmented scan with a running carry. ----
__device__ __forceinline__ float wave_incl_sum(float v) {
    const int l = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const float u = __shfl_up(v, o, 64);
        if (l >= o) v += u;
    }
    return v;
}

// Suffix (reverse inclusive) sum over lanes: lane l gets sum_{k >= l} v_k.
__device__ __forceinline__ float wave_suffix_sum(float v) {
    const int l = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const float u = __shfl_down(v, o, 64);
        if (l + o < 64) v += u;
    }
    return v;
}

__global__ void __launch_bounds__(256) k_composite_density_fwd(
    const float* __restrict__ density, const float* __restrict__ t0, const float* __restrict__ t1,
    const float* __restrict__ rgb, const int32_t* __restrict__ packed, int64_t n_rays,
    float* __restrict__ weights, float* __restrict__ trans, float* __restrict__ opacity,
    float* __restrict__ depth, float* __restrict__ comp_rgb) {
    const int64_t r = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (r >= n_rays) return;
    const int l = lane_id();
    const int start = packed[2 * r], count = packed[2 * r + 1];
    float carry = 0.0f, s_op = 0.0f, s_dep = 0.0f, s_rgb[3] = {0.0f, 0.0f, 0.0f};
    for (int c0 = 0; c0 < count; c0 += 64) {
        const bool valid = c0 + l < count;
        const int64_t j = (int64_t)start + c0 + l;
        float sd = 0.0f, m = 0.0f, col[3] = {0, 0, 0};
        if (valid) {
            const float a = t0[j], b = t1[j];
            sd = density[j] * (b - a);
            m = (a + b) * 0.5f;
            load3(rgb, j, col);
        }
        const float incl = wave_incl_sum(sd);
        float excl = __shfl_up(incl, 1, 64);
        if (l == 0) excl = 0.0f;
        const float T = expf(-(carry + excl));
        const float w = valid ? T * (1.0f - expf(-sd)) : 0.0f;
        if (valid) {
            weights[j] = w;
            trans[j] = T;
        }
        s_op += wave_sum(w);
        s_dep += wave_sum(w * m);
#pragma unroll
        for (int d = 0; d < 3; ++d) s_rgb[d] += wave_sum(w * col[d]);
        carry += __shfl(incl, 63, 64);
    }
    if (l == 0) {
        opacity[r] = s_op;
        depth[r] = s_dep;
#pragma unroll
        for (int d = 0; d < 3; ++d) comp_rgb[3 * r + d] = s_rgb[d];
    }
}

// dL/ds_i = g_i T_i exp(-s_i) - sum_{k>i} g_k w_k with g_i the total upstream
// gradient of w_i; dL/dsigma_i = dL/ds_i (t1 - t0), dL/dc_i = w_i g_rgb.
__global__ void __launch_bounds__(256) k_composite_density_bwd(
    const float* __restrict__ density, const float* __restrict__ t0, const float* __restrict__ t1,
    const float* __restrict__ rgb, const int32_t* __restrict__ packed, int64_t n_rays,
    const float* __restrict__ weights, const float* __restrict__ trans,
    const float* __restrict__ g_weights, const float* __restrict__ g_opacity,
    const float* __restrict__ g_depth, const float* __restrict__ g_comp_rgb,
    float* __restrict__ g_density, float* __restrict__ g_rgb) {
    const int64_t r = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (r >= n_rays) return;
    const int l = lane_id();
    const int start = packed[2 * r], count = packed[2 * r + 1];
    const float gop = g_opacity ? g_opacity[r] : 0.0f;
    const float gdep = g_depth ? g_depth[r] : 0.0f;
    float gcol[3] = {0, 0, 0};
    if (g_comp_rgb) load3(g_comp_rgb, r, gcol);
    float Rcarry = 0.0f;
    const int nchunks = (count + 63) / 64;
    for (int ch = nchunks - 1; ch >= 0; --ch) {
        const int c0 = ch * 64;
        const bool valid = c0 + l < count;
        const int64_t j = (int64_t)start + c0 + l;
        float gw = 0.0f, w = 0.0f, sd = 0.0f, dist = 0.0f, T = 0.0f;
        if (valid) {
            float col[3];
            load3(rgb, j, col);
            const float a = t0[j], b = t1[j];
            dist = b - a;
            sd = density[j] * dist;
            w = weights[j];
            T = trans[j];
            gw = gop + gdep * ((a + b) * 0.5f) + (gcol[0] * col[0] + gcol[1] * col[1] + gcol[2] * col[2]);
            if (g_weights) gw += g_weights[j];
        }
        const float v = gw * w;
        const float suf = wave_suffix_sum(v);
        float after = __shfl_down(suf, 1, 64);
        if (l == 63) after = 0.0f;
        const float R = Rcarry + after;
        if (valid) {
            const float gs = gw * T * expf(-sd) - R;
            g_density[j] = gs * dist;
#pragma unroll
            for (int d = 0; d < 3; ++d) g_rgb[3 * j + d] = w * gcol[d];
        }
        Rcarry += __shfl(suf, 0, 64);
    }
}

// Unbounded-sphere contraction of VolumeDensity (contract_to_unisphere,
// models/geometry.py:19-31, ContractionType.UN_BOUNDED_SPHERE), in the
// order of the reference's fp32 operations: u = (x + r) * (1 / 2r)
// (scale_anything; torch divides by a scalar as a multiply by its fp32
// reciprocal), y = 2u - 1, m = |y|; for m > 1, y = (2 - 1/m) (y / m); then
// y / 4 + 1/2.
__device__ __forceinline__ void unbounded_y(const float* __restrict__ x, int64_t i, float inv2r,
                                            float r, float y[3], float& m) {
#pragma unroll
    for (int d = 0; d < 3; ++d) y[d] = ((x[3 * i + d] + r) * inv2r) * 2.0f - 1.0f;
    m = sqrtf((y[0] * y[0] + y[1] * y[1]) + y[2] * y[2]);
}

__global__ void k_contract_unbounded(const float* __restrict__ x, int64_t n, float r, float inv2r,
                                     float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float y[3], m;
    unbounded_y(x, i, inv2r, r, y, m);
    if (m > 1.0f) {
        const float k = 2.0f - 1.0f / m;
#pragma unroll
        for (int d = 0; d < 3; ++d) y[d] = k * (y[d] / m);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) out[3 * i + d] = y[d] * 0.25f + 0.5f;
}

// Its backward: with f = phi(m) y outside the unit ball, phi = (2 - 1/m) / m,
// J^T g = phi g + (phi'(m) / m) (y . g) y and phi'(m) / m = 2 (1 - m) / m^4;
// the identity inside; then the chain factors 1/4 (output), 2 and 1/(2r).
__global__ void k_contract_unbounded_bwd(const float* __restrict__ x, const float* __restrict__ g,
                                         int64_t n, float r, float inv2r,
                                         float* __restrict__ gx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float y[3], m;
    unbounded_y(x, i, inv2r, r, y, m);
    float gy[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) gy[d] = g[3 * i + d] * 0.25f;
    if (m > 1.0f) {
        const float phi = (2.0f - 1.0f / m) / m;
        const float m2 = m * m;
        const float c = 2.0f * (1.0f - m) / (m2 * m2);
        const float yg = (y[0] * gy[0] + y[1] * gy[1]) + y[2] * gy[2];
#pragma unroll
        for (int d = 0; d < 3; ++d) gy[d] = phi * gy[d] + (c * yg) * y[d];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) gx[3 * i + d] = (gy[d] * 2.0f) * inv2r;
}

}  // namespace am

using namespace am;

static float contract_inv2r(float radius) { return 1.0f / (float)(2.0 * (double)radius); }

extern "C" int am_contract_unbounded(const float* x, int64_t n, float radius, float* out,
                                     void* stream) {
    AM_CHECK_ARG(n >= 0, "n < 0");
    AM_CHECK_ARG(radius > 0.0f, "radius %g <= 0", (double)radius);
    if (n == 0) return AM_OK;
    AM_CHECK_ARG(x && out, "NULL buffer");
    hipLaunchKernelGGL(k_contract_unbounded, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
                       to_stream(stream), x, n, radius, contract_inv2r(radius), out);
    AM_LAUNCH_CHECK();
    return AM_OK;
}

extern "C" int am_contract_unbounded_bwd(const float* x, const float* g_out, int64_t n,
                                         float radius, float* g_x, void* stream) {
    AM_CHECK_ARG(n >= 0, "n < 0");
    AM_CHECK_ARG(radius > 0.0f, "radius %g <= 0", (double)radius);
    if (n == 0) return AM_OK;
    AM_CHECK_ARG(x && g_out && g_x, "NULL buffer");
    hipLaunchKernelGGL(k_contract_unbounded_bwd, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
                       to_stream(stream), x, g_out, n, radius, contract_inv2r(radius), g_x);
    AM_LAUNCH_CHECK();
    return AM_OK;
}

extern "C" int am_sample_points(const float* rays, const int32_t* ray_indices,
                                const float* t_starts, const float* t_ends, int64_t n,
                                float* pts, float* mid, float* dists, void* stream) {
    AM_CHECK_ARG(n >= 0, "n < 0");
    if (n == 0) return AM_OK;
    AM_CHECK_ARG(rays && ray_indices && t_starts && t_ends, "NULL input");
    hipLaunchKernelGGL(k_sample_points, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
                       to_stream(stream), rays, ray_indices, t_starts, t_ends, n, pts, mid, dists);
    AM_LAUNCH_CHECK();
    return AM_OK;
}

static CompArgs comp_args(const am_composite_desc* c) {
    CompArgs a;
    a.ratio = c->cos_anneal_ratio;
    a.one_m_ratio = (float)(1.0 - (double)c->cos_anneal_ratio);
    a.kahan = c->kahan;
    return a;
}

extern "C" int am_composite_fwd(const am_composite_desc* c, const float* sdf, const float* normal,
                                const float* rays, const int32_t* ray_indices,
                                const float* t_starts, const float* t_ends, const float* rgb,
                                const float* inv_s, const int32_t* packed, int64_t n_rays,
                                int64_t n, float* weights, float* trans, float* opacity,
                                float* depth, float* comp_rgb, float* comp_normal,
                                float* comp_normal_raw, void* stream) {
    (void)ray_indices;
    AM_CHECK_ARG(c != nullptr, "composite descriptor is NULL");
    AM_CHECK_ARG(n_rays >= 0 && n >= 0, "negative size");
    if (n_rays == 0) return AM_OK;
    AM_CHECK_ARG(rays && packed && inv_s && opacity && depth && comp_rgb && comp_normal,
                 "NULL ray buffer");
    AM_CHECK_ARG(n == 0 || (sdf && normal && t_starts && t_ends && rgb && weights),
                 "NULL sample buffer");
    {
        KernelTimer kt_("composite_fwd", to_stream(stream));
        hipLaunchKernelGGL(k_composite_fwd, dim3((unsigned)ceil_div(n_rays, 4)), dim3(256), 0,
                           to_stream(stream), comp_args(c), sdf, normal, rays, t_starts, t_ends, rgb,
                           inv_s, packed, n_rays, weights, trans, opacity, depth, comp_rgb,
                           comp_normal, comp_normal_raw);
    }
    AM_LAUNCH_CHECK();
    return AM_OK;
}

extern "C" size_t am_composite_bwd_workspace_size(int64_t n_rays, int64_t n) {
    (void)n;
    WsCarver c(nullptr, 0);
    c.take<float>((size_t)(n_rays > 0 ? n_rays : 1));
    c.take<float>((size_t)colsum_tmp_floats(n_rays, 1) + 1);
    return c.used + 256;
}

extern "C" int am_composite_bwd(const am_composite_desc* c, const float* sdf, const float* normal,
                                const float* rays, const int32_t* ray_indices,
                                const float* t_starts, const float* t_ends, const float* rgb,
                                const float* inv_s, const int32_t* packed, int64_t n_rays,
                                int64_t n, const float* trans, const float* opacity,
                                const float* comp_normal_raw, const float* g_weights,
                                const float* g_opacity, const float* g_depth,
                                const float* g_comp_rgb, const float* g_comp_normal,
                                float* g_sdf, float* g_normal, float* g_rgb, float* g_inv_s,
                                void* workspace, size_t workspace_bytes, void* stream) {
    (void)ray_indices;
    AM_CHECK_ARG(c != nullptr, "composite descriptor is NULL");
    AM_CHECK_ARG(n_rays >= 0 && n >= 0, "negative size");
    AM_CHECK_ARG(g_inv_s && workspace, "NULL g_inv_s / workspace");
    hipStream_t s = to_stream(stream);
    if (n_rays == 0) return fill_f32(g_inv_s, 1, 0.0f, s);
    AM_CHECK_ARG(rays && packed && inv_s && opacity, "NULL ray buffer");
    AM_CHECK_ARG(!g_comp_normal || comp_normal_raw, "g_comp_normal needs comp_normal_raw");
    AM_CHECK_ARG(n == 0 || (sdf && normal && t_starts && t_ends && rgb && trans && g_sdf &&
                            g_normal && g_rgb),
                 "NULL sample buffer");
    WsCarver wc(workspace, workspace_bytes);
    float* per_ray = wc.take<float>((size_t)n_rays);
    float* tmp = wc.take<float>((size_t)colsum_tmp_floats(n_rays, 1) + 1);
    if (!wc.ok()) return workspace_short(wc.used, workspace_bytes);
    {
        KernelTimer kt_("composite_bwd", s);
        hipLaunchKernelGGL(k_composite_bwd, dim3((unsigned)ceil_div(n_rays, 4)), dim3(256), 0, s,
                           comp_args(c), sdf, normal, rays, t_starts, t_ends, rgb, inv_s, packed,
                           n_rays, trans, opacity, comp_normal_raw, g_weights, g_opacity, g_depth,
                           g_comp_rgb, g_comp_normal, g_sdf, g_normal, g_rgb, per_ray);
    }
    AM_LAUNCH_CHECK();
    return colsum_f32(per_ray, n_rays, 1, g_inv_s, tmp, s);
}

// ---- am_composite_bwd without a caller workspace -------------------------
// Its scratch (the per-ray partials of d loss / d inv_s and the slices of
// their fixed-order sum) lives in a file-scope device array of
// kStaticScratchBytes; each device holds its own copy, whose address
// hipGetSymbolAddress gives once per device (kept in g_scratch_dev).  Every
// partial and slice is written before it is read, so nothing is cleared;
// calls that overlap on one device share the array and race.
namespace am {

static constexpr int64_t kStaticMaxRays = int64_t(1) << 20;
static constexpr size_t kStaticScratchBytes = (size_t(4) << 20) + (size_t(80) << 10);
static constexpr int kMaxDevices = 64;

}  // namespace am

static __device__ __align__(256) unsigned char scratch[am::kStaticScratchBytes];

namespace am {

static std::mutex g_scratch_mu;
static void* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device.
static int static_scratch(void** out) {
    int dev = -1;
    AM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device %d beyond the %d the static scratch tracks", dev, kMaxDevices);
        return AM_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                      hipGetErrorString(e));
            return AM_ERR_HIP;
        }
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return AM_OK;
}

}  // namespace am

extern "C" int am_composite_bwd_noalloc(
    const am_composite_desc* c, const float* sdf, const float* normal, const float* rays,
    const int32_t* ray_indices, const float* t_starts, const float* t_ends, const float* rgb,
    const float* inv_s, const int32_t* packed, int64_t n_rays, int64_t n, const float* trans,
    const float* opacity, const float* comp_normal_raw, const float* g_weights,
    const float* g_opacity, const float* g_depth, const float* g_comp_rgb,
    const float* g_comp_normal, float* g_sdf, float* g_normal, float* g_rgb, float* g_inv_s,
    void* stream) {
    AM_CHECK_ARG(n_rays >= 0 && n >= 0, "negative size");
    const size_t need = am_composite_bwd_workspace_size(n_rays, n);
    if (n_rays > am::kStaticMaxRays || need > am::kStaticScratchBytes) {
        am::set_error("am_composite_bwd_noalloc takes at most %lld rays (got %lld): its static "
                      "scratch holds %zu bytes, the call needs %zu",
                      (long long)am::kStaticMaxRays, (long long)n_rays, am::kStaticScratchBytes,
                      need);
        return AM_ERR_INVALID;
    }
    void* ws = nullptr;
    const int rc = am::static_scratch(&ws);
    if (rc != AM_OK) return rc;
    return am_composite_bwd(c, sdf, normal, rays, ray_indices, t_starts, t_ends, rgb, inv_s,
                            packed, n_rays, n, trans, opacity, comp_normal_raw, g_weights,
                            g_opacity, g_depth, g_comp_rgb, g_comp_normal, g_sdf, g_normal, g_rgb,
                            g_inv_s, ws, am::kStaticScratchBytes, stream);
}

extern "C" int am_composite_density_fwd(const float* density, const float* t_starts,
                                        const float* t_ends, const float* rgb,
                                        const int32_t* packed, int64_t n_rays, int64_t n,
                                        float* weights, float* trans, float* opacity,
                                        float* depth, float* comp_rgb, void* stream) {
    AM_CHECK_ARG(n_rays >= 0 && n >= 0, "negative size");
    if (n_rays == 0) return AM_OK;
    AM_CHECK_ARG(packed && opacity && depth && comp_rgb, "NULL output");
    AM_CHECK_ARG(n == 0 || (density && t_starts && t_ends && rgb && weights && trans), "NULL input");
    hipStream_t s = to_stream(stream);
    {
        KernelTimer kt_("composite_bg_fwd", s);
        hipLaunchKernelGGL(k_composite_density_fwd, dim3((unsigned)ceil_div(n_rays, 4)), dim3(256),
                           0, s, density, t_starts, t_ends, rgb, packed, n_rays, weights, trans,
                           opacity, depth, comp_rgb);
    }
    AM_LAUNCH_CHECK();
    return AM_OK;
}

extern "C" int am_composite_density_bwd(const float* density, const float* t_starts,
                                        const float* t_ends, const float* rgb,
                                        const int32_t* packed, int64_t n_rays, int64_t n,
                                        const float* weights, const float* trans,
                                        const float* g_weights, const float* g_opacity,
                                        const float* g_depth, const float* g_comp_rgb,
                                        float* g_density, float* g_rgb, void* stream) {
    AM_CHECK_ARG(n_rays >= 0 && n >= 0, "negative size");
    if (n_rays == 0 || n == 0) return AM_OK;
    AM_CHECK_ARG(density && t_starts && t_ends && rgb && packed && weights && trans &&
                     g_density && g_rgb, "NULL buffer");
    hipStream_t s = to_stream(stream);
    {
        KernelTimer kt_("composite_bg_bwd", s);
        hipLaunchKernelGGL(k_composite_density_bwd, dim3((unsigned)ceil_div(n_rays, 4)), dim3(256),
                           0, s, density, t_starts, t_ends, rgb, packed, n_rays, weights, trans,
                           g_weights, g_opacity, g_depth, g_comp_rgb, g_density, g_rgb);
    }
    AM_LAUNCH_CHECK();
    return AM_OK;
}
