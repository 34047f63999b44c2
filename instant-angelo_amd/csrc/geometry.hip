// Fused VolumeSDF.forward (models/geometry.py:195-287) and
// VolumeRadiance.forward (models/texture.py:23-30) on fp32 MFMA.
//
// Geometry: one wave evaluates 16 samples.  Per sample: contraction
// (scale_anything, models/utils.py:109-114, CUDA torch semantics: divide by
// a scalar = multiply by its fp32 reciprocal), the base evaluation, 6
// central-difference evaluations (:219-234) and, with laplace, 6 around the
// tangent-shifted point (:236-275).  Each evaluation = hash-grid encode (lane
// group k gathers levels k, k+4, k+8, k+12 of its sample, producing exactly
// the MFMA B fragments of the first layer) + the SDF MLP on MFMA.  The
// backward recomputes the evaluations, back-propagates through MLP and
// encode, and emits (table row, value) pairs for the deterministic
// sort + segmented-reduce table gradient; MLP weight gradients are
// per-wave register accumulators reduced in a fixed order.
#include "hashgrid_dev.h"
#include "mlp_dev.h"
#include "table_grad.h"
#include "field_common.h"

namespace am {

using mlp::krow;

// ------------------------------------------------------------------------
// geometry
// ------------------------------------------------------------------------
struct GeoArgs {
    GridArgs g;
    float r, inv2r, eps, inv_eps, inv_pi;
    int with_grad, with_laplace, recontract, active;
    int n_in, n_out, act;
    float beta;
    int w1, b1, w2, b2, wo, bo, ld1, lds_floats;  // LDS offsets (floats)
    int64_t pw1, pb1, pw2, pb2, pwo, pbo, n_params;
    int n_evals;  // 1, 7 or 13
};

// Row-space column of the SDF MLP first layer for lane-group level slots:
// K-step s < 2LJ -> level k + 4(s/2), feature s%2; s == 2LJ -> xyz[k].
__host__ __device__ inline int geo_col_of_row(int rs, int LJ, int L) {
    const int t = rs >> 4, k = (rs & 15) >> 2, s = 4 * t + (rs & 3);
    if (s < 2 * LJ) {
        const int level = k + 4 * (s >> 1);
        return level < L ? 3 + 2 * level + (s & 1) : -1;
    }
    if (s == 2 * LJ) return k < 3 ? k : -1;
    return -1;
}

template <int LJ>
struct GeoDims {
    static constexpr int KS1 = 2 * LJ + 1;
    static constexpr int IT1 = (KS1 + 3) / 4;
    static constexpr int RS1 = 16 * IT1;
};

template <int LJ, int NH>
__device__ void geo_load_lds(const GeoArgs& a, const float* __restrict__ params, float* sm,
                             LevelParams* lv) {
    constexpr int RS1 = GeoDims<LJ>::RS1;
    for (int q = threadIdx.x; q < 64 * RS1; q += blockDim.x) {
        const int o = q / RS1, rs = q - o * RS1;
        const int col = geo_col_of_row(rs, LJ, a.g.L);
        sm[a.w1 + o * a.ld1 + rs] = col >= 0 ? params[a.pw1 + (int64_t)o * a.n_in + col] : 0.0f;
    }
    for (int q = threadIdx.x; q < 64; q += blockDim.x) sm[a.b1 + q] = params[a.pb1 + q];
    if (NH == 2) {
        for (int q = threadIdx.x; q < 64 * 64; q += blockDim.x) {
            const int o = q >> 6, c = q & 63;
            sm[a.w2 + o * 65 + c] = params[a.pw2 + q];
        }
        for (int q = threadIdx.x; q < 64; q += blockDim.x) sm[a.b2 + q] = params[a.pb2 + q];
    }
    for (int q = threadIdx.x; q < 16 * 64; q += blockDim.x) {
        const int o = q >> 6, c = q & 63;
        sm[a.wo + o * 65 + c] = o < a.n_out ? params[a.pwo + (int64_t)o * 64 + c] : 0.0f;
    }
    for (int q = threadIdx.x; q < 16; q += blockDim.x) sm[a.bo + q] = q < a.n_out ? params[a.pbo + q] : 0.0f;
    for (int q = threadIdx.x; q < a.g.L; q += blockDim.x) lv[q] = a.g.lv[q];
}

template <int LJ, int NH, bool DX>
struct GeoEval {
    float xb[GeoDims<LJ>::KS1];
    v4f z1[4];
    v4f z2[NH == 2 ? 4 : 1];
    v4f out;
    float dx[DX ? LJ : 1][2][3];
};

__device__ __forceinline__ float pick3(const float v[3], int k) {
    return k == 0 ? v[0] : (k == 1 ? v[1] : v[2]);
}

// Encode + SDF MLP of this lane's sample at contracted point x.
template <int LJ, int NH, bool DX>
__device__ __forceinline__ void geo_eval(const GeoArgs& a, const float* sm,
                                         const LevelParams* lv, const float2* __restrict__ table,
                                         const float x[3], GeoEval<LJ, NH, DX>& e) {
    constexpr int KS1 = GeoDims<LJ>::KS1;
    const int k = lane_id() >> 4;
#pragma unroll
    for (int j = 0; j < LJ; ++j) {
        const int level = k + 4 * j;
        if (level < a.active) {
            const LevelParams lp = lv[level];
            if constexpr (DX) encode_level_f2<true>(lp, x, table, &e.xb[2 * j], e.dx[j]);
            else encode_level_f2<false>(lp, x, table, &e.xb[2 * j], nullptr);
        } else {
            e.xb[2 * j] = 0.0f;
            e.xb[2 * j + 1] = 0.0f;
            if constexpr (DX) {
#pragma unroll
                for (int f = 0; f < 2; ++f)
#pragma unroll
                    for (int d = 0; d < 3; ++d) e.dx[j][f][d] = 0.0f;
            }
        }
    }
    e.xb[2 * LJ] = k < 3 ? pick3(x, k) * 2.0f - 1.0f : 0.0f;
#pragma unroll
    for (int t = 0; t < 4; ++t) e.z1[t] = mlp::bias_tile(sm + a.b1, t, 64);
    mlp::mm_fwd<KS1, 4>(sm + a.w1, a.ld1, e.xb, e.z1);
    v4f h[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) h[t] = mlp::act4(a.act, e.z1[t], a.beta);
    if constexpr (NH == 2) {
#pragma unroll
        for (int t = 0; t < 4; ++t) e.z2[t] = mlp::bias_tile(sm + a.b2, t, 64);
        mlp::mm_fwd<16, 4>(sm + a.w2, 65, reinterpret_cast<const float*>(h), e.z2);
#pragma unroll
        for (int t = 0; t < 4; ++t) h[t] = mlp::act4(a.act, e.z2[t], a.beta);
    }
    e.out = mlp::bias_tile(sm + a.bo, 0, a.n_out);
    mlp::mm_fwd<16, 1>(sm + a.wo, 65, reinterpret_cast<const float*>(h), &e.out);
}

__device__ __forceinline__ float contract1(float v, const GeoArgs& a) { return (v + a.r) * a.inv2r; }

__device__ __forceinline__ float clampr(float v, float r) { return fminf(fmaxf(v, -r), r); }

// FD point j (axis j/2, sign + for even j) around world point P, contracted.
__device__ __forceinline__ void fd_point(const GeoArgs& a, const float P[3], int j, float x[3],
                                         bool mask[3]) {
    const int ax = j >> 1;
    const float off = (j & 1) ? -a.eps : a.eps;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float v = P[d] + (d == ax ? off : 0.0f);
        mask[d] = v >= -a.r && v <= a.r;
        x[d] = contract1(clampr(v, a.r), a);
    }
}

__device__ __forceinline__ void normalize3(const float v[3], float out[3], float* nrm) {
    const float n = sqrtf((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    const float d = fmaxf(n, 1e-12f);
#pragma unroll
    for (int q = 0; q < 3; ++q) out[q] = v[q] / d;
    if (nrm) *nrm = n;
}

// Backward of y = v / max(|v|, 1e-12) (F.normalize).
__device__ __forceinline__ void normalize3_bwd(const float v[3], const float g[3], float out[3]) {
    const float n = sqrtf((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    const float d = fmaxf(n, 1e-12f);
    const float gv = g[0] * v[0] + g[1] * v[1] + g[2] * v[2];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        float r = g[q] / d;
        if (n >= 1e-12f) r -= gv / (d * d) * (v[q] / n);
        out[q] = r;
    }
}

__device__ __forceinline__ void cross3(const float a[3], const float b[3], float c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// Everything the forward computes per sample (all lanes of the sample).
struct GeoSample {
    float P[3], x0[3];
    float f[6], grad[3], n[3];
    float rn[3], ps[3], fs[6], grad_s[3], ns[3], dot;
};

template <int LJ, int NH>
__device__ __forceinline__ void geo_shifted_point(const GeoArgs& a, const GeoSample& s, int j,
                                                  float x[3], bool mask[3]) {
    // recontract: the shifted point is in [0,1] coordinates and the FD
    // offsets / clamp / contraction are applied to it again (reference).
    // Otherwise the shift is applied in world space.
    const int ax = j >> 1;
    const float off = (j & 1) ? -a.eps : a.eps;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float base = s.ps[d];
        const float v = base + (d == ax ? off : 0.0f);
        mask[d] = v >= -a.r && v <= a.r;
        x[d] = contract1(clampr(v, a.r), a);
    }
}

// Forward part shared by fwd and bwd: base output, FD, tangent shift, laplace.
template <int LJ, int NH>
__device__ __forceinline__ void geo_sample_fwd(const GeoArgs& a, const float* sm,
                                               const LevelParams* lv, const float2* table,
                                               const float* __restrict__ pts,
                                               const float* __restrict__ randn, int64_t n,
                                               bool valid, GeoSample& s, v4f* base_out) {
    const int p = lane_id() & 15;
#pragma unroll
    for (int d = 0; d < 3; ++d) s.P[d] = valid ? pts[3 * n + d] : 0.0f;
#pragma unroll
    for (int d = 0; d < 3; ++d) s.x0[d] = contract1(s.P[d], a);
    {
        GeoEval<LJ, NH, false> e;
        geo_eval<LJ, NH, false>(a, sm, lv, table, s.x0, e);
        if (base_out) *base_out = e.out;
    }
    if (!a.with_grad) return;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        float x[3];
        bool m[3];
        fd_point(a, s.P, j, x, m);
        GeoEval<LJ, NH, false> e;
        geo_eval<LJ, NH, false>(a, sm, lv, table, x, e);
        s.f[j] = shfl(e.out[0], p);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) s.grad[d] = (0.5f * (s.f[2 * d] - s.f[2 * d + 1])) * a.inv_eps;
    normalize3(s.grad, s.n, nullptr);
    if (!a.with_laplace) return;
    float rd[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) rd[d] = valid ? randn[3 * n + d] : 1.0f;
    normalize3(rd, s.rn, nullptr);
    float t[3];
    cross3(s.n, s.rn, t);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float base = a.recontract ? s.x0[d] : s.P[d];
        s.ps[d] = base + t[d] * a.eps;
    }
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        float x[3];
        bool m[3];
        geo_shifted_point<LJ, NH>(a, s, j, x, m);
        GeoEval<LJ, NH, false> e;
        geo_eval<LJ, NH, false>(a, sm, lv, table, x, e);
        s.fs[j] = shfl(e.out[0], p);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) s.grad_s[d] = (0.5f * (s.fs[2 * d] - s.fs[2 * d + 1])) * a.inv_eps;
    normalize3(s.grad_s, s.ns, nullptr);
    s.dot = (s.n[0] * s.ns[0] + s.n[1] * s.ns[1]) + s.n[2] * s.ns[2];
}

template <int LJ, int NH>
__global__ void __launch_bounds__(256) k_geo_fwd(GeoArgs a, const float* __restrict__ pts,
                                                 const float* __restrict__ randn, int64_t N,
                                                 const float* __restrict__ table_f,
                                                 const float* __restrict__ params,
                                                 float* __restrict__ sdf, float* __restrict__ grad,
                                                 float* __restrict__ normal,
                                                 float* __restrict__ feature,
                                                 float* __restrict__ laplace) {
    extern __shared__ float sm[];
    __shared__ LevelParams lv[AM_MAX_LEVELS];
    geo_load_lds<LJ, NH>(a, params, sm, lv);
    __syncthreads();
    const float2* table = reinterpret_cast<const float2*>(table_f);
    const int l = lane_id(), p = l & 15, k = l >> 4;
    const int nw = blockDim.x >> 6, wave = threadIdx.x >> 6;
    const int64_t n_tiles = (N + 15) / 16;
    const int fdim = a.n_out + 3;
    for (int64_t tile = (int64_t)blockIdx.x * nw + wave; tile < n_tiles;
         tile += (int64_t)gridDim.x * nw) {
        const int64_t n = tile * 16 + p;
        const bool valid = n < N;
        GeoSample s;
        v4f out;
        geo_sample_fwd<LJ, NH>(a, sm, lv, table, pts, randn, n, valid, s, &out);
        if (!valid) continue;
        if (k == 0) sdf[n] = out[0];
        if (feature) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 4 * k + r;
                if (row < a.n_out) feature[n * fdim + row] = out[r];
            }
            if (k < 3) feature[n * fdim + a.n_out + k] = pick3(s.x0, k) * 2.0f - 1.0f;
        }
        if (k == 0) {
            if (grad && a.with_grad)
#pragma unroll
                for (int d = 0; d < 3; ++d) grad[3 * n + d] = s.grad[d];
            if (normal && a.with_grad)
#pragma unroll
                for (int d = 0; d < 3; ++d) normal[3 * n + d] = s.n[d];
            if (laplace && a.with_laplace) {
                const float lo = -1.0f + 1e-6f, hi = 1.0f - 1e-6f;
                const float dc = fminf(fmaxf(s.dot, lo), hi);
                laplace[n] = acosf(dc) * a.inv_pi;
            }
        }
    }
}

// ---- backward ----
template <int LJ, int NH>
struct GeoGrad {
    v4f dW1[4 * GeoDims<LJ>::IT1];
    v4f dW2[NH == 2 ? 16 : 1];
    v4f dWo[4];
    v4f db1[4];
    v4f db2[NH == 2 ? 4 : 1];
    v4f dbo;
};

struct EmitCtx {
    uint32_t* keys;
    Vals<2>* vals;
    uint32_t sentinel;
};

// MLP backward of one evaluation with output gradient `go` (D tile rows =
// outputs); accumulates weight gradients, emits table pairs and, with DX,
// returns the contracted-position gradient (summed over the 4 lane groups).
template <int LJ, int NH, bool DX>
__device__ __forceinline__ void geo_eval_bwd(const GeoArgs& a, const float* sm,
                                             const LevelParams* lv, const float x[3],
                                             const GeoEval<LJ, NH, DX>& e, v4f go,
                                             float* scratch, GeoGrad<LJ, NH>& G,
                                             const EmitCtx& em, int64_t pair_base, bool valid,
                                             float gpos[3]) {
    constexpr int KS1 = GeoDims<LJ>::KS1, IT1 = GeoDims<LJ>::IT1;
    const int l = lane_id(), k = l >> 4;
    v4f h[4];
    const v4f* zl = NH == 2 ? e.z2 : e.z1;
#pragma unroll
    for (int t = 0; t < 4; ++t) h[t] = mlp::act4(a.act, zl[t], a.beta);
    mlp::mm_dw<1, 4, 16>(&go, reinterpret_cast<const float*>(h), scratch, G.dWo);
    G.dbo += go;
    v4f gh[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) gh[t] = v4f{0, 0, 0, 0};
    mlp::mm_bwd_in<4, 4>(sm + a.wo, 65, reinterpret_cast<const float*>(&go), gh);
    if constexpr (NH == 2) {
        v4f gz2[4], h1[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
#pragma unroll
            for (int r = 0; r < 4; ++r) gz2[t][r] = gh[t][r] * mlp::act_bwd(a.act, e.z2[t][r], a.beta);
            h1[t] = mlp::act4(a.act, e.z1[t], a.beta);
            G.db2[t] += gz2[t];
        }
        mlp::mm_dw<4, 4, 16>(gz2, reinterpret_cast<const float*>(h1), scratch, G.dW2);
#pragma unroll
        for (int t = 0; t < 4; ++t) gh[t] = v4f{0, 0, 0, 0};
        mlp::mm_bwd_in<16, 4>(sm + a.w2, 65, reinterpret_cast<const float*>(gz2), gh);
    }
    v4f gz1[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
        for (int r = 0; r < 4; ++r) gz1[t][r] = gh[t][r] * mlp::act_bwd(a.act, e.z1[t][r], a.beta);
        G.db1[t] += gz1[t];
    }
    mlp::mm_dw<4, IT1, KS1>(gz1, e.xb, scratch, G.dW1);
    v4f gin[IT1];
#pragma unroll
    for (int t = 0; t < IT1; ++t) gin[t] = v4f{0, 0, 0, 0};
    mlp::mm_bwd_in<16, IT1>(sm + a.w1, a.ld1, reinterpret_cast<const float*>(gz1), gin);
    const float* gb = reinterpret_cast<const float*>(gin);  // gb[s] = grad of B fragment s
    // table pairs: slot ((j*8 + c) * 64 + lane) after pair_base
#pragma unroll
    for (int j = 0; j < LJ; ++j) {
        const int level = k + 4 * j;
        const bool on = valid && level < a.active;
        LevelPos pp{};
        LevelParams lp{};
        if (on) {
            lp = lv[level];
            pp = level_pos(lp, x);
        }
        const float g0 = gb[2 * j], g1 = gb[2 * j + 1];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int64_t slot = pair_base + (int64_t)(j * 8 + c) * 64 + l;
            if (on) {
                const float w = corner_weight(pp, c);
                em.keys[slot] = corner_row(lp, pp, c);
                Vals<2> v;
                v.v[0] = w * g0;
                v.v[1] = w * g1;
                em.vals[slot] = v;
            } else {
                em.keys[slot] = em.sentinel;
            }
        }
    }
    if constexpr (DX) {
        float gp[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int j = 0; j < LJ; ++j)
#pragma unroll
            for (int d = 0; d < 3; ++d)
                gp[d] += gb[2 * j] * e.dx[j][0][d] + gb[2 * j + 1] * e.dx[j][1][d];
        if (k < 3) {
            const float gxyz = gb[2 * LJ] * 2.0f;
#pragma unroll
            for (int d = 0; d < 3; ++d) gp[d] += d == k ? gxyz : 0.0f;
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) gpos[d] = sum_groups(gp[d]);
    }
}

// Row-0-only output gradient tile (FD / curvature evaluations).
__device__ __forceinline__ v4f go_row0(float g) {
    const int k = lane_id() >> 4;
    return v4f{k == 0 ? g : 0.0f, 0.0f, 0.0f, 0.0f};
}

template <int LJ, int NH>
__global__ void __launch_bounds__(256) k_geo_bwd(
    GeoArgs a, const float* __restrict__ pts, const float* __restrict__ randn, int64_t N,
    const float* __restrict__ table_f, const float* __restrict__ params,
    const float* __restrict__ g_sdf, const float* __restrict__ g_grad,
    const float* __restrict__ g_normal, const float* __restrict__ g_feature,
    const float* __restrict__ g_laplace, EmitCtx em, float* __restrict__ partial) {
    extern __shared__ float sm[];
    __shared__ LevelParams lv[AM_MAX_LEVELS];
    geo_load_lds<LJ, NH>(a, params, sm, lv);
    __syncthreads();
    const float2* table = reinterpret_cast<const float2*>(table_f);
    const int l = lane_id(), p = l & 15, k = l >> 4;
    const int nw = blockDim.x >> 6, wave = threadIdx.x >> 6;
    float* scratch = sm + a.lds_floats + wave * (128 * mlp::kScratchLd);
    const int64_t n_tiles = (N + 15) / 16;
    const int fdim = a.n_out + 3;
    const int64_t pairs_per_eval = (int64_t)LJ * 8 * 64;
    GeoGrad<LJ, NH> G;
#pragma unroll
    for (int q = 0; q < 4 * GeoDims<LJ>::IT1; ++q) G.dW1[q] = v4f{0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < (NH == 2 ? 16 : 1); ++q) G.dW2[q] = v4f{0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        G.dWo[q] = v4f{0, 0, 0, 0};
        G.db1[q] = v4f{0, 0, 0, 0};
    }
#pragma unroll
    for (int q = 0; q < (NH == 2 ? 4 : 1); ++q) G.db2[q] = v4f{0, 0, 0, 0};
    G.dbo = v4f{0, 0, 0, 0};

    for (int64_t tile = (int64_t)blockIdx.x * nw + wave; tile < n_tiles;
         tile += (int64_t)gridDim.x * nw) {
        const int64_t n = tile * 16 + p;
        const bool valid = n < N;
        const int64_t pair_tile = tile * a.n_evals * pairs_per_eval;
        GeoSample s;
        geo_sample_fwd<LJ, NH>(a, sm, lv, table, pts, randn, n, valid, s, nullptr);
        // upstream gradients of this sample
        float gg[3] = {0, 0, 0}, gnrm[3] = {0, 0, 0};
        if (valid && g_grad)
#pragma unroll
            for (int d = 0; d < 3; ++d) gg[d] = g_grad[3 * n + d];
        if (valid && g_normal)
#pragma unroll
            for (int d = 0; d < 3; ++d) gnrm[d] = g_normal[3 * n + d];
        int ev = 0;
        if (a.with_laplace) {
            const float gl = (valid && g_laplace) ? g_laplace[n] : 0.0f;
            const float lo = -1.0f + 1e-6f, hi = 1.0f - 1e-6f;
            const float dc = fminf(fmaxf(s.dot, lo), hi);
            const bool pass = s.dot >= lo && s.dot <= hi;
            const float g_dot = pass ? (gl * a.inv_pi) * (-1.0f / sqrtf(1.0f - dc * dc)) : 0.0f;
            float g_ns[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                gnrm[d] += g_dot * s.ns[d];
                g_ns[d] = g_dot * s.n[d];
            }
            float g_gs[3];
            normalize3_bwd(s.grad_s, g_ns, g_gs);
            float g_ps[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int j = 0; j < 6; ++j) {
                const float gf = (j & 1 ? -0.5f : 0.5f) * a.inv_eps * g_gs[j >> 1];
                float x[3];
                bool m[3];
                geo_shifted_point<LJ, NH>(a, s, j, x, m);
                GeoEval<LJ, NH, true> e;
                geo_eval<LJ, NH, true>(a, sm, lv, table, x, e);
                float gpos[3];
                geo_eval_bwd<LJ, NH, true>(a, sm, lv, x, e, go_row0(gf), scratch, G, em,
                                           pair_tile + (int64_t)(ev++) * pairs_per_eval, valid,
                                           gpos);
#pragma unroll
                for (int d = 0; d < 3; ++d)
                    if (m[d]) g_ps[d] += gpos[d] * a.inv2r;
            }
            float g_t[3], gcr[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) g_t[d] = g_ps[d] * a.eps;
            cross3(s.rn, g_t, gcr);
#pragma unroll
            for (int d = 0; d < 3; ++d) gnrm[d] += gcr[d];
        }
        if (a.with_grad) {
            float gn2g[3];
            normalize3_bwd(s.grad, gnrm, gn2g);
#pragma unroll
            for (int d = 0; d < 3; ++d) gg[d] += gn2g[d];
#pragma unroll
            for (int j = 0; j < 6; ++j) {
                const float gf = (j & 1 ? -0.5f : 0.5f) * a.inv_eps * gg[j >> 1];
                float x[3];
                bool m[3];
                fd_point(a, s.P, j, x, m);
                GeoEval<LJ, NH, false> e;
                geo_eval<LJ, NH, false>(a, sm, lv, table, x, e);
                geo_eval_bwd<LJ, NH, false>(a, sm, lv, x, e, go_row0(gf), scratch, G, em,
                                            pair_tile + (int64_t)(ev++) * pairs_per_eval, valid,
                                            nullptr);
            }
        }
        {
            v4f go;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 4 * k + r;
                float v = 0.0f;
                if (valid && row < a.n_out) {
                    if (g_feature) v = g_feature[n * fdim + row];
                    if (row == 0 && g_sdf) v += g_sdf[n];
                }
                go[r] = v;
            }
            GeoEval<LJ, NH, false> e;
            geo_eval<LJ, NH, false>(a, sm, lv, table, s.x0, e);
            geo_eval_bwd<LJ, NH, false>(a, sm, lv, s.x0, e, go, scratch, G, em,
                                        pair_tile + (int64_t)(ev++) * pairs_per_eval, valid,
                                        nullptr);
        }
    }
    // per-wave partial gradient row in parameter layout
    float* out = partial + ((int64_t)blockIdx.x * nw + wave) * a.n_params;
    const int i = p;
    constexpr int IT1 = GeoDims<LJ>::IT1;
#pragma unroll
    for (int ot = 0; ot < 4; ++ot)
#pragma unroll
        for (int it = 0; it < IT1; ++it)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int o = 16 * ot + 4 * k + r;
                const int col = geo_col_of_row(16 * it + i, LJ, a.g.L);
                if (col >= 0) out[a.pw1 + (int64_t)o * a.n_in + col] = G.dW1[ot * IT1 + it][r];
            }
    if constexpr (NH == 2) {
#pragma unroll
        for (int ot = 0; ot < 4; ++ot)
#pragma unroll
            for (int it = 0; it < 4; ++it)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    out[a.pw2 + (16 * ot + 4 * k + r) * 64 + 16 * it + i] = G.dW2[ot * 4 + it][r];
    }
#pragma unroll
    for (int it = 0; it < 4; ++it)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int o = 4 * k + r;
            if (o < a.n_out) out[a.pwo + o * 64 + 16 * it + i] = G.dWo[it][r];
        }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const v4f s1 = mlp::rowsum_tile(G.db1[t]);
        if (i == 0)
#pragma unroll
            for (int r = 0; r < 4; ++r) out[a.pb1 + 16 * t + 4 * k + r] = s1[r];
        if constexpr (NH == 2) {
            const v4f s2 = mlp::rowsum_tile(G.db2[t]);
            if (i == 0)
#pragma unroll
                for (int r = 0; r < 4; ++r) out[a.pb2 + 16 * t + 4 * k + r] = s2[r];
        }
    }
    const v4f so = mlp::rowsum_tile(G.dbo);
    if (i == 0)
#pragma unroll
        for (int r = 0; r < 4; ++r)
            if (4 * k + r < a.n_out) out[a.pbo + 4 * k + r] = so[r];
}

// ------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------
static int make_geo_args(const am_grid_desc* g, const am_mlp_desc* m, const am_sdf_desc* s,
                         GeoArgs* out, int* LJ) {
    AM_CHECK_ARG(g && m && s, "NULL descriptor");
    if (int rc = check_mlp_common(m)) return rc;
    AM_CHECK_ARG(g->n_features == 2, "fused geometry needs n_features = 2 (got %d)", g->n_features);
    AM_CHECK_ARG(g->n_levels >= 1 && g->n_levels <= 16, "fused geometry needs <= 16 levels");
    AM_CHECK_ARG(m->n_in == 3 + 2 * g->n_levels, "SDF MLP n_in %d != 3 + L*F", m->n_in);
    AM_CHECK_ARG(s->active_levels >= 0 && s->active_levels <= g->n_levels, "active_levels %d",
                 s->active_levels);
    AM_CHECK_ARG(s->radius > 0.0f, "radius <= 0");
    AM_CHECK_ARG(!s->with_grad || s->fd_eps > 0.0f, "fd_eps <= 0");
    AM_CHECK_ARG(!s->with_laplace || s->with_grad, "laplace needs with_grad");
    GeoArgs a{};
    a.g = make_grid_args(g);
    a.r = s->radius;
    a.inv2r = 1.0f / (float)(2.0 * (double)s->radius);
    a.eps = s->fd_eps;
    a.inv_eps = 1.0f / s->fd_eps;
    a.inv_pi = 1.0f / (float)M_PI;
    a.with_grad = s->with_grad;
    a.with_laplace = s->with_laplace;
    a.recontract = s->recontract;
    a.active = s->active_levels;
    a.n_in = m->n_in;
    a.n_out = m->n_out;
    a.act = m->hidden_act;
    a.beta = m->softplus_beta;
    // level slots per lane group: 2 (L <= 8) or 4 (L <= 16); spare slots
    // hold levels >= L and stay zero.
    const int lj = g->n_levels <= 8 ? 2 : 4;
    const int rs1 = 16 * ((2 * lj + 1 + 3) / 4);
    a.ld1 = rs1 + 1;
    int o = 0;
    a.w1 = o; o += 64 * a.ld1;
    a.b1 = o; o += 64;
    a.w2 = o; a.b2 = o;
    if (m->n_hidden_layers == 2) { a.w2 = o; o += 64 * 65; a.b2 = o; o += 64; }
    a.wo = o; o += 16 * 65;
    a.bo = o; o += 16;
    a.lds_floats = o;
    layout_params(m->n_in, m->n_out, m->n_hidden_layers, &a.pw1, &a.pb1, &a.pw2, &a.pb2, &a.pwo,
                  &a.pbo, &a.n_params);
    a.n_evals = 1 + (s->with_grad ? 6 : 0) + (s->with_laplace ? 6 : 0);
    *out = a;
    *LJ = lj;
    return AM_OK;
}

template <int LJ, int NH>
static int geo_fwd_launch(const GeoArgs& a, const float* pts, const float* randn, int64_t n,
                          const float* table, const float* params, float* sdf, float* grad,
                          float* normal, float* feature, float* laplace, hipStream_t s) {
    const size_t lds = (size_t)a.lds_floats * sizeof(float);
    AM_HIP(hipFuncSetAttribute((const void*)k_geo_fwd<LJ, NH>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    {
        KernelTimer kt_("geo_fwd", s);
        hipLaunchKernelGGL((k_geo_fwd<LJ, NH>), dim3(field_blocks(n)), dim3(256), lds, s, a, pts, randn,
                           n, table, params, sdf, grad, normal, feature, laplace);
    }
    AM_LAUNCH_CHECK();
    return AM_OK;
}

static int64_t geo_pairs(const GeoArgs& a, int LJ, int64_t n) {
    return ((n + 15) / 16) * (int64_t)a.n_evals * LJ * 8 * 64;
}

template <int LJ, int NH>
static int geo_bwd_launch(const GeoArgs& a, const am_grid_desc* g, const float* pts,
                          const float* randn, int64_t n, const float* table, const float* params,
                          const float* g_sdf, const float* g_grad, const float* g_normal,
                          const float* g_feature, const float* g_laplace, float* dtable,
                          float* dparams, void* ws, size_t ws_bytes, hipStream_t s) {
    const int blocks = field_blocks(n);
    const int waves = blocks * 4;
    WsCarver c(ws, ws_bytes);
    float* partial = c.take<float>((size_t)waves * a.n_params);
    float* ctmp = c.take<float>((size_t)colsum_tmp_floats(waves, a.n_params) + 1);
    TableReduceWs tw;
    const int64_t m = geo_pairs(a, LJ, n);
    if (int rc = table_reduce_carve(c, 2, m, g->n_entries, &tw)) return rc;
    if (!c.ok()) {
        set_error("workspace too small: need %zu, got %zu", c.used, ws_bytes);
        return AM_ERR_WORKSPACE;
    }
    EmitCtx em{tw.keys[0], reinterpret_cast<Vals<2>*>(tw.vals[0]), table_sentinel(tw.end_bit)};
    const size_t lds = (size_t)(a.lds_floats + 4 * 128 * mlp::kScratchLd) * sizeof(float);
    AM_HIP(hipFuncSetAttribute((const void*)k_geo_bwd<LJ, NH>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    {
        KernelTimer kt_("geo_bwd", s);
        hipLaunchKernelGGL((k_geo_bwd<LJ, NH>), dim3(blocks), dim3(256), lds, s, a, pts, randn, n, table,
                           params, g_sdf, g_grad, g_normal, g_feature, g_laplace, em, partial);
    }
    AM_LAUNCH_CHECK();
    if (int rc = colsum_f32(partial, waves, a.n_params, dparams, ctmp, s)) return rc;
    return table_reduce_run(tw, 2, g->n_entries, dtable, s);
}

#define AM_GEO_DISPATCH(FN, ...)                                                   \
    do {                                                                           \
        const bool two = m->n_hidden_layers == 2;                                  \
        switch (LJ) {                                                              \
            case 2: return two ? FN<2, 2>(__VA_ARGS__) : FN<2, 1>(__VA_ARGS__);    \
            default: return two ? FN<4, 2>(__VA_ARGS__) : FN<4, 1>(__VA_ARGS__);   \
        }                                                                          \
    } while (0)

}  // namespace am

using namespace am;

extern "C" int am_sdf_fwd(const am_grid_desc* g, const am_mlp_desc* m, const am_sdf_desc* sd,
                          const float* pts, const float* randn, int64_t n, const float* table,
                          const float* params, float* sdf, float* grad, float* normal,
                          float* feature, float* laplace, void* stream) {
    GeoArgs a;
    int LJ;
    if (int rc = make_geo_args(g, m, sd, &a, &LJ)) return rc;
    AM_CHECK_ARG(n >= 0, "n < 0");
    if (n == 0) return AM_OK;
    AM_CHECK_ARG(pts && table && params && sdf, "NULL buffer");
    AM_CHECK_ARG(!sd->with_laplace || randn, "laplace needs randn");
    hipStream_t s = to_stream(stream);
    AM_GEO_DISPATCH(geo_fwd_launch, a, pts, randn, n, table, params, sdf, grad, normal, feature,
                    laplace, s);
}

extern "C" size_t am_sdf_bwd_workspace_size(const am_grid_desc* g, const am_mlp_desc* m,
                                            const am_sdf_desc* sd, int64_t n) {
    GeoArgs a;
    int LJ;
    if (make_geo_args(g, m, sd, &a, &LJ) || n < 0) return 0;
    const int waves = field_blocks(n) * 4;
    WsCarver c(nullptr, 0);
    c.take<float>((size_t)waves * a.n_params);
    c.take<float>((size_t)colsum_tmp_floats(waves, a.n_params) + 1);
    TableReduceWs tw;
    if (table_reduce_carve(c, 2, geo_pairs(a, LJ, n), g->n_entries, &tw)) return 0;
    return c.used + 256;
}

extern "C" int am_sdf_bwd(const am_grid_desc* g, const am_mlp_desc* m, const am_sdf_desc* sd,
                          const float* pts, const float* randn, int64_t n, const float* table,
                          const float* params, const float* g_sdf, const float* g_grad,
                          const float* g_normal, const float* g_feature, const float* g_laplace,
                          float* dtable, float* dparams, void* workspace, size_t workspace_bytes,
                          void* stream) {
    GeoArgs a;
    int LJ;
    if (int rc = make_geo_args(g, m, sd, &a, &LJ)) return rc;
    AM_CHECK_ARG(n >= 0, "n < 0");
    AM_CHECK_ARG(table && params && dtable && dparams && workspace, "NULL buffer");
    AM_CHECK_ARG(n == 0 || pts, "NULL pts");
    AM_CHECK_ARG(!sd->with_laplace || n == 0 || randn, "laplace needs randn");
    hipStream_t s = to_stream(stream);
    AM_GEO_DISPATCH(geo_bwd_launch, a, g, pts, randn, n, table, params, g_sdf, g_grad, g_normal,
                    g_feature, g_laplace, dtable, dparams, workspace, workspace_bytes, s);
}

