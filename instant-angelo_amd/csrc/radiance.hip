// Fused VolumeRadiance.forward (models/texture.py:23-30) on fp32 MFMA: the
// SH encoding of the view direction (tcnn SphericalHarmonics, :17,24-25) is
// computed in the kernel prologue, cat(feature, sh, normal) feeds the
// radiance MLP and the colour activation (sigmoid) is applied in-register.
// Backward returns d/d feature, d/d normal and deterministic weight grads.
#include "mlp_dev.h"
#include "field_common.h"

namespace am {

struct RadArgs {
    int n_in, n_out, act, fdim, sh_deg, n_sh;
    int base_color;  // VolumeDualColor: + sigmoid(feature[:, 1:4])
    float beta;
    int w1, b1, w2, b2, wo, bo, ld1, lds_floats;
    int64_t pw1, pb1, pw2, pb2, pwo, pbo, n_params;
};

// Rows of a wave's transpose scratch for mm_dw: the 4 output tiles of a
// hidden layer plus the first layer's IT1 = ceil(KS/4) input tiles.
__host__ __device__ constexpr int rad_scratch_rows(int KS) {
    return 16 * (4 + (KS + 3) / 4) > 128 ? 16 * (4 + (KS + 3) / 4) : 128;
}

// Radiance first-layer input column c sits at K-step c/4, lane group c%4,
// i.e. row krow(c/4, c%4) of the row space.
__device__ __forceinline__ int rad_col_of_row(int rs) {
    const int t = rs >> 4, k = (rs & 15) >> 2, s = 4 * t + (rs & 3);
    return 4 * s + k;
}

// Real SH bases at v (degree <= 4 bands), tcnn / svox2 ordering and signs.
__device__ __forceinline__ void sh_eval_v(const float v[3], int deg, float* sh) {
    const float x = v[0], y = v[1], z = v[2];
    const float xx = x * x, yy = y * y, zz = z * z;
    // normalisations sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!) (x sqrt2 for m != 0)
    const float C0 = 0.28209479177387814f;   // 1/(2 sqrt(pi))
    const float C1 = 0.48860251190291992f;   // sqrt(3)/(2 sqrt(pi))
    const float C2a = 1.0925484305920792f;   // sqrt(15)/(2 sqrt(pi))
    const float C2b = 0.31539156525252005f;  // sqrt(5)/(4 sqrt(pi))
    const float C2c = 0.54627421529603959f;  // sqrt(15)/(4 sqrt(pi))
    const float C3a = 0.59004358992664352f;  // sqrt(35/2)/(4 sqrt(pi))
    const float C3b = 2.8906114426405538f;   // sqrt(105)/(2 sqrt(pi))
    const float C3c = 0.45704579946446572f;  // sqrt(21/2)/(4 sqrt(pi))
    const float C3d = 0.37317633259011540f;  // sqrt(7)/(4 sqrt(pi))
    const float C3e = 1.4453057213202769f;   // sqrt(105)/(4 sqrt(pi))
    sh[0] = C0;
    if (deg > 1) {
        sh[1] = -C1 * y;
        sh[2] = C1 * z;
        sh[3] = -C1 * x;
    }
    if (deg > 2) {
        sh[4] = C2a * x * y;
        sh[5] = -C2a * y * z;
        sh[6] = C2b * (3.0f * zz - 1.0f);
        sh[7] = -C2a * x * z;
        sh[8] = C2c * (xx - yy);
    }
    if (deg > 3) {
        sh[9] = -C3a * y * (3.0f * xx - yy);
        sh[10] = C3b * x * y * z;
        sh[11] = -C3c * y * (5.0f * zz - 1.0f);
        sh[12] = C3d * z * (5.0f * zz - 3.0f);
        sh[13] = -C3c * x * (5.0f * zz - 1.0f);
        sh[14] = C3e * z * (xx - yy);
        sh[15] = -C3a * x * (xx - 3.0f * yy);
    }
}

// VolumeRadiance maps the unit direction d to (d+1)/2 (models/texture.py:24)
// and tcnn maps its [0,1] input back by 2*x - 1; the round trip is kept.
__device__ __forceinline__ void sh_eval_dir(const float d[3], int deg, float* sh) {
    float v[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const float d01 = (d[q] + 1.0f) * 0.5f;
        v[q] = d01 * 2.0f - 1.0f;
    }
    sh_eval_v(v, deg, sh);
}

template <int NH>
__device__ void rad_load_lds(const RadArgs& a, const float* __restrict__ params, float* sm,
                             int RS1) {
    for (int q = threadIdx.x; q < 64 * RS1; q += blockDim.x) {
        const int o = q / RS1, rs = q - o * RS1;
        const int col = rad_col_of_row(rs);
        sm[a.w1 + o * a.ld1 + rs] = col < a.n_in ? params[a.pw1 + (int64_t)o * a.n_in + col] : 0.0f;
    }
    for (int q = threadIdx.x; q < 64; q += blockDim.x) sm[a.b1 + q] = params[a.pb1 + q];
    if (NH == 2) {
        for (int q = threadIdx.x; q < 64 * 64; q += blockDim.x)
            sm[a.w2 + (q >> 6) * 65 + (q & 63)] = params[a.pw2 + q];
        for (int q = threadIdx.x; q < 64; q += blockDim.x) sm[a.b2 + q] = params[a.pb2 + q];
    }
    for (int q = threadIdx.x; q < 16 * 64; q += blockDim.x) {
        const int o = q >> 6, c = q & 63;
        sm[a.wo + o * 65 + c] = o < a.n_out ? params[a.pwo + (int64_t)o * 64 + c] : 0.0f;
    }
    for (int q = threadIdx.x; q < 16; q += blockDim.x) sm[a.bo + q] = q < a.n_out ? params[a.pbo + q] : 0.0f;
}

template <int KS, int NH>
struct RadEval {
    float xb[KS];
    v4f z1[4];
    v4f z2[NH == 2 ? 4 : 1];
    v4f out;
};

template <int KS, int NH, int PR>
__device__ __forceinline__ void rad_eval(const RadArgs& a, const float* sm,
                                         const float* __restrict__ feature,
                                         const float* __restrict__ rays,
                                         const int32_t* __restrict__ ridx,
                                         const float* __restrict__ normal, int64_t n, bool valid,
                                         RadEval<KS, NH>& e) {
    const int k = lane_id() >> 4;
    float sh[16];
    float d[3] = {0.0f, 0.0f, 1.0f};
    if (valid) {
        const int64_t r = ridx[n];
#pragma unroll
        for (int q = 0; q < 3; ++q) d[q] = rays[6 * r + 3 + q];
    }
    sh_eval_dir(d, a.sh_deg, sh);
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int c = 4 * s + k;
        float v = 0.0f;
        if (valid && c < a.n_in) {
            if (c < a.fdim) {
                v = feature[n * a.fdim + c];
            } else if (c < a.fdim + a.n_sh) {
                const int q = c - a.fdim;
                float t = sh[0];
#pragma unroll
                for (int u = 1; u < 16; ++u) t = (q == u) ? sh[u] : t;
                v = t;
            } else {
                v = normal[3 * n + (c - a.fdim - a.n_sh)];
            }
        }
        e.xb[s] = v;
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) e.z1[t] = mlp::bias_tile(sm + a.b1, t, 64);
    mlp::mm_fwd<KS, 4, mlp::FragF, PR>(sm + a.w1, a.ld1, mlp::FragF{e.xb}, e.z1);
    v4f h[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) h[t] = mlp::act4(a.act, e.z1[t], a.beta);
    if constexpr (NH == 2) {
#pragma unroll
        for (int t = 0; t < 4; ++t) e.z2[t] = mlp::bias_tile(sm + a.b2, t, 64);
        mlp::mm_fwd<16, 4, mlp::FragV, PR>(sm + a.w2, 65, mlp::FragV{h}, e.z2);
#pragma unroll
        for (int t = 0; t < 4; ++t) h[t] = mlp::act4(a.act, e.z2[t], a.beta);
    }
    e.out = mlp::bias_tile(sm + a.bo, 0, a.n_out);
    mlp::mm_fwd<16, 1, mlp::FragV, PR>(sm + a.wo, 65, mlp::FragV{h}, &e.out);
}

template <int KS, int NH, int PR>
__global__ void __launch_bounds__(256) k_rad_fwd(RadArgs a, const float* __restrict__ feature,
                                                 const float* __restrict__ rays,
                                                 const int32_t* __restrict__ ridx,
                                                 const float* __restrict__ normal, int64_t N,
                                                 const float* __restrict__ params,
                                                 float* __restrict__ rgb) {
    extern __shared__ float sm[];
    rad_load_lds<NH>(a, params, sm, 16 * ((KS + 3) / 4));
    __syncthreads();
    const int l = lane_id(), p = l & 15, k = l >> 4;
    const int nw = blockDim.x >> 6, wave = threadIdx.x >> 6;
    const int64_t n_tiles = (N + 15) / 16;
    for (int64_t tile = (int64_t)blockIdx.x * nw + wave; tile < n_tiles;
         tile += (int64_t)gridDim.x * nw) {
        const int64_t n = tile * 16 + p;
        const bool valid = n < N;
        RadEval<KS, NH> e;
        rad_eval<KS, NH, PR>(a, sm, feature, rays, ridx, normal, n, valid, e);
        if (valid && k == 0)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                if (r < a.n_out) {
                    float c = 1.0f / (1.0f + expf(-e.out[r]));
                    if (a.base_color) c += 1.0f / (1.0f + expf(-feature[n * a.fdim + 1 + r]));
                    rgb[n * a.n_out + r] = c;
                }
    }
}

template <int KS, int NH, int PR>
__global__ void __launch_bounds__(256) k_rad_bwd(RadArgs a, const float* __restrict__ feature,
                                                 const float* __restrict__ rays,
                                                 const int32_t* __restrict__ ridx,
                                                 const float* __restrict__ normal, int64_t N,
                                                 const float* __restrict__ params,
                                                 const float* __restrict__ g_rgb,
                                                 float* __restrict__ g_feature,
                                                 float* __restrict__ g_normal,
                                                 float* __restrict__ partial) {
    constexpr int IT1 = (KS + 3) / 4;
    extern __shared__ float sm[];
    rad_load_lds<NH>(a, params, sm, 16 * IT1);
    __syncthreads();
    const int l = lane_id(), p = l & 15, k = l >> 4;
    const int nw = blockDim.x >> 6, wave = threadIdx.x >> 6;
    float* scratch = sm + a.lds_floats + wave * (rad_scratch_rows(KS) * mlp::kScratchLd);
    const int64_t n_tiles = (N + 15) / 16;
    v4f dW1[4 * IT1], dW2[NH == 2 ? 16 : 1], dWo[4], db1[4], db2[NH == 2 ? 4 : 1], dbo;
#pragma unroll
    for (int q = 0; q < 4 * IT1; ++q) dW1[q] = v4f{0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < (NH == 2 ? 16 : 1); ++q) dW2[q] = v4f{0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 4; ++q) { dWo[q] = v4f{0, 0, 0, 0}; db1[q] = v4f{0, 0, 0, 0}; }
#pragma unroll
    for (int q = 0; q < (NH == 2 ? 4 : 1); ++q) db2[q] = v4f{0, 0, 0, 0};
    dbo = v4f{0, 0, 0, 0};
    for (int64_t tile = (int64_t)blockIdx.x * nw + wave; tile < n_tiles;
         tile += (int64_t)gridDim.x * nw) {
        const int64_t n = tile * 16 + p;
        const bool valid = n < N;
        RadEval<KS, NH> e;
        rad_eval<KS, NH, PR>(a, sm, feature, rays, ridx, normal, n, valid, e);
        v4f go;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 4 * k + r;
            float v = 0.0f;
            if (valid && row < a.n_out) {
                const float y = 1.0f / (1.0f + expf(-e.out[r]));
                v = g_rgb[n * a.n_out + row] * ((1.0f - y) * y);
            }
            go[r] = v;
        }
        v4f h[4];
        const v4f* zl = NH == 2 ? e.z2 : e.z1;
#pragma unroll
        for (int t = 0; t < 4; ++t) h[t] = mlp::act4(a.act, zl[t], a.beta);
        mlp::mm_dw<1, 4, 16, mlp::FragV, PR>(&go, mlp::FragV{h}, scratch, dWo);
        dbo += go;
        v4f gh[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) gh[t] = v4f{0, 0, 0, 0};
        mlp::mm_bwd_in<4, 4, mlp::FragV, PR>(sm + a.wo, 65, mlp::FragV{&go}, gh);
        if constexpr (NH == 2) {
            v4f gz2[4], h1[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
#pragma unroll
                for (int r = 0; r < 4; ++r) gz2[t][r] = gh[t][r] * mlp::act_bwd(a.act, e.z2[t][r], a.beta);
                h1[t] = mlp::act4(a.act, e.z1[t], a.beta);
                db2[t] += gz2[t];
            }
            mlp::mm_dw<4, 4, 16, mlp::FragV, PR>(gz2, mlp::FragV{h1}, scratch, dW2);
#pragma unroll
            for (int t = 0; t < 4; ++t) gh[t] = v4f{0, 0, 0, 0};
            mlp::mm_bwd_in<16, 4, mlp::FragV, PR>(sm + a.w2, 65, mlp::FragV{gz2}, gh);
        }
        v4f gz1[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
#pragma unroll
            for (int r = 0; r < 4; ++r) gz1[t][r] = gh[t][r] * mlp::act_bwd(a.act, e.z1[t][r], a.beta);
            db1[t] += gz1[t];
        }
        mlp::mm_dw<4, IT1, KS, mlp::FragF, PR>(gz1, mlp::FragF{e.xb}, scratch, dW1);
        v4f gin[IT1];
#pragma unroll
        for (int t = 0; t < IT1; ++t) gin[t] = v4f{0, 0, 0, 0};
        mlp::mm_bwd_in<16, IT1, mlp::FragV, PR>(sm + a.w1, a.ld1, mlp::FragV{gz1}, gin);
        if (valid) {
#pragma unroll
            for (int t = 0; t < IT1; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int s = 4 * t + r;
                    const int c = 4 * s + k;
                    if (s >= KS || c >= a.n_in) continue;
                    if (c < a.fdim) {
                        float gv = gin[t][r];
                        if (a.base_color && c >= 1 && c <= 3) {   // d sigmoid(feature[c]) term
                            const float y = 1.0f / (1.0f + expf(-feature[n * a.fdim + c]));
                            gv += g_rgb[n * a.n_out + c - 1] * ((1.0f - y) * y);
                        }
                        if (g_feature) g_feature[n * a.fdim + c] = gv;
                    } else if (c >= a.fdim + a.n_sh) {
                        if (g_normal) g_normal[3 * n + (c - a.fdim - a.n_sh)] = gin[t][r];
                    }
                }
        }
    }
    float* out = partial + ((int64_t)blockIdx.x * nw + wave) * a.n_params;
    const int i = p;
#pragma unroll
    for (int ot = 0; ot < 4; ++ot)
#pragma unroll
        for (int it = 0; it < IT1; ++it)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int o = 16 * ot + 4 * k + r;
                const int col = rad_col_of_row(16 * it + i);
                if (col < a.n_in) out[a.pw1 + (int64_t)o * a.n_in + col] = dW1[ot * IT1 + it][r];
            }
    if constexpr (NH == 2) {
#pragma unroll
        for (int ot = 0; ot < 4; ++ot)
#pragma unroll
            for (int it = 0; it < 4; ++it)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    out[a.pw2 + (16 * ot + 4 * k + r) * 64 + 16 * it + i] = dW2[ot * 4 + it][r];
    }
#pragma unroll
    for (int it = 0; it < 4; ++it)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int o = 4 * k + r;
            if (o < a.n_out) out[a.pwo + o * 64 + 16 * it + i] = dWo[it][r];
        }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const v4f s1 = mlp::rowsum_tile(db1[t]);
        if (i == 0)
#pragma unroll
            for (int r = 0; r < 4; ++r) out[a.pb1 + 16 * t + 4 * k + r] = s1[r];
        if constexpr (NH == 2) {
            const v4f s2 = mlp::rowsum_tile(db2[t]);
            if (i == 0)
#pragma unroll
                for (int r = 0; r < 4; ++r) out[a.pb2 + 16 * t + 4 * k + r] = s2[r];
        }
    }
    const v4f so = mlp::rowsum_tile(dbo);
    if (i == 0)
#pragma unroll
        for (int r = 0; r < 4; ++r)
            if (4 * k + r < a.n_out) out[a.pbo + 4 * k + r] = so[r];
}

__global__ void k_sh(const float* __restrict__ d01, int64_t n, int deg, float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // tcnn's input convention: directions in [0,1], mapped by 2*x - 1.
    const float v[3] = {d01[3 * i] * 2.0f - 1.0f, d01[3 * i + 1] * 2.0f - 1.0f,
                        d01[3 * i + 2] * 2.0f - 1.0f};
    float sh[16];
    sh_eval_v(v, deg, sh);
    const int m = deg * deg;
    for (int q = 0; q < m; ++q) out[i * m + q] = sh[q];
}

// host side
static int make_rad_args(const am_mlp_desc* m, int sh_degree, int base_color, int fdim,
                         RadArgs* out, int* KS) {
    if (int rc = check_mlp_common(m)) return rc;
    AM_CHECK_ARG(base_color == 0 || base_color == 1, "base_color %d", base_color);
    AM_CHECK_ARG(!base_color || (fdim >= 4 && m->n_out == 3),
                 "base colour needs feature_dim >= 4 and 3 outputs");
    AM_CHECK_ARG(sh_degree >= 1 && sh_degree <= 4, "sh_degree %d out of [1,4]", sh_degree);
    AM_CHECK_ARG(fdim >= 0 && m->n_in == fdim + sh_degree * sh_degree + 3,
                 "radiance n_in %d != feature_dim + degree^2 + 3", m->n_in);
    AM_CHECK_ARG(m->n_in <= 96, "radiance n_in %d > 96", m->n_in);
    RadArgs a{};
    a.n_in = m->n_in;
    a.n_out = m->n_out;
    a.act = m->hidden_act;
    a.beta = m->softplus_beta;
    a.fdim = fdim;
    a.base_color = base_color;
    a.sh_deg = sh_degree;
    a.n_sh = sh_degree * sh_degree;
    const int ks = (m->n_in + 3) / 4;
    // K-step counts compiled: 8, 10, 12 (inputs <= 48) and 24 (<= 96, the
    // Neuralangelo texture's 65 + 3 features, 16 SH and the normal)
    const int ks_t = ks <= 8 ? 8 : (ks <= 10 ? 10 : (ks <= 12 ? 12 : 24));
    const int rs1 = 16 * ((ks_t + 3) / 4);
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
    *out = a;
    *KS = ks_t;
    return AM_OK;
}

template <int KS, int NH, int PR>
static int rad_fwd_launch(const RadArgs& a, const float* feature, const float* rays,
                          const int32_t* ridx, const float* normal, int64_t n,
                          const float* params, float* rgb, hipStream_t s) {
    const size_t lds = (size_t)a.lds_floats * sizeof(float);
    AM_HIP(hipFuncSetAttribute((const void*)k_rad_fwd<KS, NH, PR>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    {
        KernelTimer kt_("rad_fwd", s);
        // one round of resident blocks (its registers allow 3 waves a SIMD)
        const int blocks = persistent_blocks(
            (n + 15) / 16, 4, resident_blocks((const void*)k_rad_fwd<KS, NH, PR>, 256, lds));
        hipLaunchKernelGGL((k_rad_fwd<KS, NH, PR>), dim3(blocks), dim3(256), lds, s, a, feature,
                           rays, ridx, normal, n, params, rgb);
    }
    AM_LAUNCH_CHECK();
    return AM_OK;
}

template <int KS, int NH, int PR>
static int rad_bwd_launch(const RadArgs& a, const float* feature, const float* rays,
                          const int32_t* ridx, const float* normal, int64_t n,
                          const float* params, const float* g_rgb, float* g_feature,
                          float* g_normal, float* dparams, void* ws, size_t ws_bytes,
                          hipStream_t s) {
    const int blocks = bwd_blocks(n);
    const int waves = blocks * 4;
    WsCarver c(ws, ws_bytes);
    float* partial = c.take<float>((size_t)waves * a.n_params);
    float* ctmp = c.take<float>((size_t)colsum_tmp_floats(waves, a.n_params) + 1);
    if (!c.ok()) return workspace_short(c.used, ws_bytes);
    const size_t lds =
        (size_t)(a.lds_floats + 4 * rad_scratch_rows(KS) * mlp::kScratchLd) * sizeof(float);
    AM_CHECK_ARG(lds <= 160 * 1024, "radiance backward needs %zu bytes of LDS", lds);
    AM_HIP(hipFuncSetAttribute((const void*)k_rad_bwd<KS, NH, PR>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    {
        KernelTimer kt_("rad_bwd", s);
        hipLaunchKernelGGL((k_rad_bwd<KS, NH, PR>), dim3(blocks), dim3(256), lds, s, a, feature, rays, ridx,
                           normal, n, params, g_rgb, g_feature, g_normal, partial);
    }
    AM_LAUNCH_CHECK();
    return colsum_f32(partial, waves, a.n_params, dparams, ctmp, s);
}

#define AM_RAD_DISPATCH_PR(FN, PR, ...)                                                   \
    do {                                                                                  \
        const bool two = m->n_hidden_layers == 2;                                         \
        switch (KS) {                                                                     \
            case 8: return two ? FN<8, 2, PR>(__VA_ARGS__) : FN<8, 1, PR>(__VA_ARGS__);    \
            case 10: return two ? FN<10, 2, PR>(__VA_ARGS__) : FN<10, 1, PR>(__VA_ARGS__); \
            case 12: return two ? FN<12, 2, PR>(__VA_ARGS__) : FN<12, 1, PR>(__VA_ARGS__); \
            default: return two ? FN<24, 2, PR>(__VA_ARGS__) : FN<24, 1, PR>(__VA_ARGS__); \
        }                                                                                 \
    } while (0)

#define AM_RAD_DISPATCH(FN, ...)                                      \
    do {                                                              \
        if (m->precision == AM_PREC_MIXED)                            \
            AM_RAD_DISPATCH_PR(FN, mlp::kBF16, __VA_ARGS__);          \
        AM_RAD_DISPATCH_PR(FN, mlp::kF32, __VA_ARGS__);               \
    } while (0)

}  // namespace am

using namespace am;

extern "C" int am_radiance_fwd(const am_mlp_desc* m, int32_t sh_degree, int32_t base_color,
                               const float* feature, int32_t feature_dim, const float* rays,
                               const int32_t* ray_indices, const float* normal, int64_t n,
                               const float* params, float* rgb, void* stream) {
    RadArgs a;
    int KS;
    if (int rc = make_rad_args(m, sh_degree, base_color, feature_dim, &a, &KS)) return rc;
    AM_CHECK_ARG(n >= 0, "n < 0");
    if (n == 0) return AM_OK;
    AM_CHECK_ARG(feature && rays && ray_indices && normal && params && rgb, "NULL buffer");
    hipStream_t s = to_stream(stream);
    AM_RAD_DISPATCH(rad_fwd_launch, a, feature, rays, ray_indices, normal, n, params, rgb, s);
}

extern "C" size_t am_radiance_bwd_workspace_size(const am_mlp_desc* m, int64_t n) {
    if (!m || n < 0) return 0;
    const int waves = bwd_blocks(n) * 4;
    int64_t pw1, pb1, pw2, pb2, pwo, pbo, np;
    layout_params(m->n_in, m->n_out, m->n_hidden_layers, &pw1, &pb1, &pw2, &pb2, &pwo, &pbo, &np);
    WsCarver c(nullptr, 0);
    c.take<float>((size_t)waves * np);
    c.take<float>((size_t)colsum_tmp_floats(waves, np) + 1);
    return c.used + 256;
}

extern "C" int am_radiance_bwd(const am_mlp_desc* m, int32_t sh_degree, int32_t base_color,
                               const float* feature, int32_t feature_dim, const float* rays,
                               const int32_t* ray_indices, const float* normal, int64_t n,
                               const float* params, const float* g_rgb, float* g_feature,
                               float* g_normal, float* dparams, void* workspace,
                               size_t workspace_bytes, void* stream) {
    RadArgs a;
    int KS;
    if (int rc = make_rad_args(m, sh_degree, base_color, feature_dim, &a, &KS)) return rc;
    AM_CHECK_ARG(n >= 0, "n < 0");
    AM_CHECK_ARG(params && dparams && workspace, "NULL buffer");
    AM_CHECK_ARG(n == 0 || (feature && rays && ray_indices && normal && g_rgb), "NULL input");
    hipStream_t s = to_stream(stream);
    AM_RAD_DISPATCH(rad_bwd_launch, a, feature, rays, ray_indices, normal, n, params, g_rgb,
                    g_feature, g_normal, dparams, workspace, workspace_bytes, s);
}

extern "C" int am_sh_fwd(const float* dirs01, int64_t n, int32_t degree, float* out, void* stream) {
    AM_CHECK_ARG(degree >= 1 && degree <= 4, "degree %d out of [1,4]", degree);
    AM_CHECK_ARG(n >= 0, "n < 0");
    if (n == 0) return AM_OK;
    AM_CHECK_ARG(dirs01 && out, "NULL buffer");
    hipLaunchKernelGGL(k_sh, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, to_stream(stream),
                       dirs01, n, degree, out);
    AM_LAUNCH_CHECK();
    return AM_OK;
}
