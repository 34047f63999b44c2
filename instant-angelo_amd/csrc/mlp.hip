// Generic VanillaMLP forward / backward on fp32 MFMA (get_mlp,
// models/network_utils.py:96-140,177-185).  The fused field kernels share
// the engine in mlp_dev.h; these entry points serve direct get_mlp users.
#include "mlp_dev.h"

namespace am {

struct MlpArgs {
    int n_in, n_out, n_layers, act, out_act;
    float beta;
    int w_lds[5], b_lds[5], ld[5], rows_out[5], n_rows_lds[5], cols_in[5];
    int64_t w_par[5], b_par[5];
    int lds_floats;  // weights + biases
    int64_t n_params;
};

static int make_mlp_args(const am_mlp_desc* m, MlpArgs* a) {
    AM_CHECK_ARG(m != nullptr, "mlp descriptor is NULL");
    AM_CHECK_ARG(m->n_neurons == 64, "n_neurons %d: only 64 is implemented", m->n_neurons);
    AM_CHECK_ARG(m->n_in >= 1 && m->n_in <= 64, "n_in %d out of [1,64]", m->n_in);
    if (m->precision != AM_PREC_FP32) {
        set_error("the stand-alone MLP runs in fp32 (precision %d)", m->precision);
        return AM_ERR_UNSUPPORTED;
    }
    AM_CHECK_ARG(m->n_out >= 1 && m->n_out <= 64, "n_out %d out of [1,64]", m->n_out);
    AM_CHECK_ARG(m->n_hidden_layers >= 1 && m->n_hidden_layers <= 4, "n_hidden_layers %d",
                 m->n_hidden_layers);
    AM_CHECK_ARG(m->hidden_act == AM_ACT_RELU || m->hidden_act == AM_ACT_SOFTPLUS ||
                     m->hidden_act == AM_ACT_NONE,
                 "hidden_act %d", m->hidden_act);
    AM_CHECK_ARG(m->output_act >= AM_ACT_NONE && m->output_act <= AM_ACT_SIGMOID, "output_act %d",
                 m->output_act);
    MlpArgs r{};
    r.n_in = m->n_in;
    r.n_out = m->n_out;
    r.n_layers = m->n_hidden_layers + 1;
    r.act = m->hidden_act;
    r.out_act = m->output_act;
    r.beta = m->softplus_beta;
    int lds = 0;
    int64_t par = 0;
    for (int L = 0; L < r.n_layers; ++L) {
        const int in = L == 0 ? m->n_in : 64;
        const int out = L == r.n_layers - 1 ? m->n_out : 64;
        const int in_rows = 16 * (int)ceil_div(in, 16);
        r.cols_in[L] = in;
        r.rows_out[L] = out;
        r.n_rows_lds[L] = 16 * (int)ceil_div(out, 16);
        r.ld[L] = in_rows + 1;
        r.w_lds[L] = lds;
        lds += r.n_rows_lds[L] * r.ld[L];
        r.b_lds[L] = lds;
        lds += r.n_rows_lds[L];
        r.w_par[L] = par;
        par += (int64_t)out * in;
        r.b_par[L] = par;
        par += out;
    }
    r.lds_floats = lds;
    r.n_params = par;
    *a = r;
    return AM_OK;
}

// Weights and biases -> LDS (zero padding outside the real rows / columns).
__device__ void load_mlp_lds(const MlpArgs& a, const float* __restrict__ params, float* sm) {
    for (int L = 0; L < a.n_layers; ++L) {
        const int rows = a.n_rows_lds[L], ld = a.ld[L];
        for (int q = threadIdx.x; q < rows * ld; q += blockDim.x) {
            const int o = q / ld, c = q - o * ld;
            sm[a.w_lds[L] + q] = (o < a.rows_out[L] && c < a.cols_in[L])
                                     ? params[a.w_par[L] + (int64_t)o * a.cols_in[L] + c]
                                     : 0.0f;
        }
        for (int q = threadIdx.x; q < rows; q += blockDim.x)
            sm[a.b_lds[L] + q] = q < a.rows_out[L] ? params[a.b_par[L] + q] : 0.0f;
    }
}

template <int OT, class Frag>
__device__ __forceinline__ void out_layer(const MlpArgs& a, const float* sm, Frag h, v4f* o) {
    const int L = a.n_layers - 1;
#pragma unroll
    for (int t = 0; t < OT; ++t) o[t] = mlp::bias_tile(sm + a.b_lds[L], t, a.rows_out[L]);
    mlp::mm_fwd<16, OT>(sm + a.w_lds[L], a.ld[L], h, o);
}

// Forward of one 16-point tile; keeps pre-activations for backward.
template <int KS, int OT>
__device__ __forceinline__ void tile_fwd(const MlpArgs& a, const float* sm, const float* xb,
                                         v4f (&z)[4][4], v4f* zo) {
    {
        v4f acc[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t] = mlp::bias_tile(sm + a.b_lds[0], t, 64);
        mlp::mm_fwd<KS, 4>(sm + a.w_lds[0], a.ld[0], mlp::FragF{xb}, acc);
#pragma unroll
        for (int t = 0; t < 4; ++t) z[0][t] = acc[t];
    }
#pragma unroll
    for (int L = 1; L < 4; ++L) {
        if (L < a.n_layers - 1) {
            v4f h[4], acc[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                h[t] = mlp::act4(a.act, z[L - 1][t], a.beta);
                acc[t] = mlp::bias_tile(sm + a.b_lds[L], t, 64);
            }
            mlp::mm_fwd<16, 4>(sm + a.w_lds[L], a.ld[L], mlp::FragV{h}, acc);
#pragma unroll
            for (int t = 0; t < 4; ++t) z[L][t] = acc[t];
        }
    }
    v4f h[4];
    const int Lh = a.n_layers - 2;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        v4f zz = z[0][t];
#pragma unroll
        for (int L = 1; L < 4; ++L)
            if (L == Lh) zz = z[L][t];
        h[t] = mlp::act4(a.act, zz, a.beta);
    }
    out_layer<OT>(a, sm, mlp::FragV{h}, zo);
}

template <int KS>
__device__ __forceinline__ void load_x(const MlpArgs& a, const float* __restrict__ x, int64_t p,
                                       bool valid, float* xb) {
    const int k = lane_id() >> 4;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int c = mlp::krow(s, k);
        xb[s] = (valid && c < a.n_in) ? x[p * a.n_in + c] : 0.0f;
    }
}

template <int KS, int OT>
__global__ void __launch_bounds__(256) k_mlp_fwd(MlpArgs a, const float* __restrict__ x, int64_t n,
                                                 const float* __restrict__ params,
                                                 float* __restrict__ y) {
    extern __shared__ float sm[];
    load_mlp_lds(a, params, sm);
    __syncthreads();
    const int l = lane_id(), i = l & 15, k = l >> 4;
    const int nw = blockDim.x >> 6, wave = threadIdx.x >> 6;
    const int64_t n_tiles = (n + 15) / 16;
    for (int64_t tile = (int64_t)blockIdx.x * nw + wave; tile < n_tiles;
         tile += (int64_t)gridDim.x * nw) {
        const int64_t p = tile * 16 + i;
        const bool valid = p < n;
        float xb[KS];
        load_x<KS>(a, x, p, valid, xb);
        v4f z[4][4], zo[OT];
        tile_fwd<KS, OT>(a, sm, xb, z, zo);
        if (valid) {
#pragma unroll
            for (int t = 0; t < OT; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * t + 4 * k + r;
                    if (row < a.n_out) y[p * a.n_out + row] = mlp::act_fwd(a.out_act, zo[t][r], a.beta);
                }
        }
    }
}

// Adds a register dW tile set (OT x IT tiles) into the LDS accumulator.
template <int OT, int IT>
__device__ __forceinline__ void add_dw_lds(float* acc, int ld, const v4f* dW) {
    const int l = lane_id(), i = l & 15, k = l >> 4;
#pragma unroll
    for (int ot = 0; ot < OT; ++ot)
#pragma unroll
        for (int it = 0; it < IT; ++it)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[(16 * ot + 4 * k + r) * ld + 16 * it + i] += dW[ot * IT + it][r];
}

template <int T>
__device__ __forceinline__ void add_db_lds(float* acc, const v4f* g) {
    const int l = lane_id(), i = l & 15, k = l >> 4;
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const v4f s = mlp::rowsum_tile(g[t]);
        if (i == 0)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[16 * t + 4 * k + r] += s[r];
    }
}

// One wave per block; LDS = weights | dW accumulator (same layout) | scratch.
template <int KS, int OT>
__global__ void __launch_bounds__(64) k_mlp_bwd(MlpArgs a, const float* __restrict__ x, int64_t n,
                                                const float* __restrict__ params,
                                                const float* __restrict__ dy,
                                                float* __restrict__ dx,
                                                float* __restrict__ partial) {
    extern __shared__ float sm[];
    float* gacc = sm + a.lds_floats;
    float* scratch = gacc + a.lds_floats;
    load_mlp_lds(a, params, sm);
    for (int q = threadIdx.x; q < a.lds_floats; q += blockDim.x) gacc[q] = 0.0f;
    __syncthreads();
    const int l = lane_id(), i = l & 15, k = l >> 4;
    const int64_t n_tiles = (n + 15) / 16;
    const int Lo = a.n_layers - 1;
    for (int64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const int64_t p = tile * 16 + i;
        const bool valid = p < n;
        float xb[KS];
        load_x<KS>(a, x, p, valid, xb);
        v4f z[4][4], zo[OT];
        tile_fwd<KS, OT>(a, sm, xb, z, zo);
        v4f go[OT];
#pragma unroll
        for (int t = 0; t < OT; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * t + 4 * k + r;
                go[t][r] = (valid && row < a.n_out)
                               ? dy[p * a.n_out + row] * mlp::act_bwd(a.out_act, zo[t][r], a.beta)
                               : 0.0f;
            }
        // output layer
        v4f gh[4];
        {
            v4f hin[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                v4f zz = z[0][t];
#pragma unroll
                for (int L = 1; L < 4; ++L)
                    if (L == Lo - 1) zz = z[L][t];
                hin[t] = mlp::act4(a.act, zz, a.beta);
            }
            v4f dW[OT * 4];
#pragma unroll
            for (int q = 0; q < OT * 4; ++q) dW[q] = v4f{0, 0, 0, 0};
            mlp::mm_dw<OT, 4, 16>(go, mlp::FragV{hin}, scratch, dW);
            add_dw_lds<OT, 4>(gacc + a.w_lds[Lo], a.ld[Lo], dW);
            add_db_lds<OT>(gacc + a.b_lds[Lo], go);
#pragma unroll
            for (int t = 0; t < 4; ++t) gh[t] = v4f{0, 0, 0, 0};
            mlp::mm_bwd_in<4 * OT, 4>(sm + a.w_lds[Lo], a.ld[Lo], mlp::FragV{go}, gh);
        }
        // hidden layers, last to first
#pragma unroll
        for (int L = 3; L >= 0; --L) {
            if (L >= Lo) continue;
            v4f gz[4];
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) gz[t][r] = gh[t][r] * mlp::act_bwd(a.act, z[L][t][r], a.beta);
            add_db_lds<4>(gacc + a.b_lds[L], gz);
            if (L == 0) {
                v4f dW[4 * (KS / 4)];
#pragma unroll
                for (int q = 0; q < 4 * (KS / 4); ++q) dW[q] = v4f{0, 0, 0, 0};
                mlp::mm_dw<4, KS / 4, KS>(gz, mlp::FragF{xb}, scratch, dW);
                add_dw_lds<4, KS / 4>(gacc + a.w_lds[0], a.ld[0], dW);
                if (dx) {
                    v4f gx[KS / 4];
#pragma unroll
                    for (int t = 0; t < KS / 4; ++t) gx[t] = v4f{0, 0, 0, 0};
                    mlp::mm_bwd_in<16, KS / 4>(sm + a.w_lds[0], a.ld[0], mlp::FragV{gz}, gx);
                    if (valid)
#pragma unroll
                        for (int t = 0; t < KS / 4; ++t)
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const int row = 16 * t + 4 * k + r;
                                if (row < a.n_in) dx[p * a.n_in + row] = gx[t][r];
                            }
                }
            } else {
                v4f hin[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) hin[t] = mlp::act4(a.act, z[L - 1][t], a.beta);
                v4f dW[16];
#pragma unroll
                for (int q = 0; q < 16; ++q) dW[q] = v4f{0, 0, 0, 0};
                mlp::mm_dw<4, 4, 16>(gz, mlp::FragV{hin}, scratch, dW);
                add_dw_lds<4, 4>(gacc + a.w_lds[L], a.ld[L], dW);
#pragma unroll
                for (int t = 0; t < 4; ++t) gh[t] = v4f{0, 0, 0, 0};
                mlp::mm_bwd_in<16, 4>(sm + a.w_lds[L], a.ld[L], mlp::FragV{gz}, gh);
            }
        }
    }
    __syncthreads();
    // LDS accumulator -> this block's partial row in parameter layout.
    float* out = partial + (int64_t)blockIdx.x * a.n_params;
    for (int L = 0; L < a.n_layers; ++L) {
        const int in = a.cols_in[L], ld = a.ld[L];
        for (int q = threadIdx.x; q < a.rows_out[L] * in; q += blockDim.x) {
            const int o = q / in, c = q - o * in;
            out[a.w_par[L] + q] = gacc[a.w_lds[L] + o * ld + c];
        }
        for (int q = threadIdx.x; q < a.rows_out[L]; q += blockDim.x)
            out[a.b_par[L] + q] = gacc[a.b_lds[L] + q];
    }
}

static constexpr int kMlpBwdMaxBlocks = 1024;

static size_t mlp_bwd_lds_bytes(const MlpArgs& a) {
    return (size_t)(2 * a.lds_floats + (16 * 4 + 16 * 4) * mlp::kScratchLd) * sizeof(float);
}

template <int KS, int OT>
static int launch_fwd(const MlpArgs& a, const float* x, int64_t n, const float* params, float* y,
                      hipStream_t s) {
    const size_t lds = (size_t)a.lds_floats * sizeof(float);
    AM_HIP(hipFuncSetAttribute((const void*)k_mlp_fwd<KS, OT>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const int blocks = persistent_blocks((n + 15) / 16, 4, 2048);
    hipLaunchKernelGGL((k_mlp_fwd<KS, OT>), dim3(blocks), dim3(256), lds, s, a, x, n, params, y);
    AM_LAUNCH_CHECK();
    return AM_OK;
}

template <int KS, int OT>
static int launch_bwd(const MlpArgs& a, const float* x, int64_t n, const float* params,
                      const float* dy, float* dx, float* dparams, void* ws, size_t ws_bytes,
                      hipStream_t s) {
    const size_t lds = mlp_bwd_lds_bytes(a);
    AM_CHECK_ARG(lds <= 160 * 1024, "MLP too large for the LDS-resident backward (%zu bytes)", lds);
    const int blocks = persistent_blocks((n + 15) / 16, 1, kMlpBwdMaxBlocks);
    WsCarver c(ws, ws_bytes);
    float* partial = c.take<float>((size_t)blocks * a.n_params);
    float* tmp = c.take<float>((size_t)colsum_tmp_floats(blocks, a.n_params) + 1);
    if (!c.ok()) return workspace_short(c.used, ws_bytes);
    AM_HIP(hipFuncSetAttribute((const void*)k_mlp_bwd<KS, OT>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_mlp_bwd<KS, OT>), dim3(blocks), dim3(64), lds, s, a, x, n, params, dy, dx,
                       partial);
    AM_LAUNCH_CHECK();
    return colsum_f32(partial, blocks, a.n_params, dparams, tmp, s);
}

#define AM_MLP_DISPATCH(FN, ...)                                              \
    do {                                                                      \
        const int ks = 4 * (int)ceil_div(a.n_in, 16);                         \
        const bool big = a.n_out > 16;                                        \
        switch (ks) {                                                         \
            case 4: return big ? FN<4, 4>(__VA_ARGS__) : FN<4, 1>(__VA_ARGS__);     \
            case 8: return big ? FN<8, 4>(__VA_ARGS__) : FN<8, 1>(__VA_ARGS__);     \
            case 12: return big ? FN<12, 4>(__VA_ARGS__) : FN<12, 1>(__VA_ARGS__);  \
            default: return big ? FN<16, 4>(__VA_ARGS__) : FN<16, 1>(__VA_ARGS__);  \
        }                                                                     \
    } while (0)

}  // namespace am

using namespace am;

extern "C" int64_t am_mlp_n_params(const am_mlp_desc* m) {
    MlpArgs a;
    if (make_mlp_args(m, &a)) return -1;
    return a.n_params;
}

extern "C" int am_mlp_fwd(const am_mlp_desc* m, const float* x, int64_t n, const float* params,
                          float* y, void* stream) {
    MlpArgs a;
    if (int rc = make_mlp_args(m, &a)) return rc;
    AM_CHECK_ARG(n >= 0, "n < 0");
    if (n == 0) return AM_OK;
    AM_CHECK_ARG(x && params && y, "NULL buffer");
    hipStream_t s = to_stream(stream);
    AM_MLP_DISPATCH(launch_fwd, a, x, n, params, y, s);
}

extern "C" size_t am_mlp_bwd_workspace_size(const am_mlp_desc* m, int64_t n) {
    MlpArgs a;
    if (make_mlp_args(m, &a) || n < 0) return 0;
    const int blocks = persistent_blocks((n + 15) / 16, 1, kMlpBwdMaxBlocks);
    WsCarver c(nullptr, 0);
    c.take<float>((size_t)blocks * a.n_params);
    c.take<float>((size_t)colsum_tmp_floats(blocks, a.n_params) + 1);
    return c.used + 256;
}

extern "C" int am_mlp_bwd(const am_mlp_desc* m, const float* x, int64_t n, const float* params,
                          const float* dy, float* dx, float* dparams, void* workspace,
                          size_t workspace_bytes, void* stream) {
    MlpArgs a;
    if (int rc = make_mlp_args(m, &a)) return rc;
    AM_CHECK_ARG(n >= 0, "n < 0");
    AM_CHECK_ARG(params && dparams, "NULL params/dparams");
    AM_CHECK_ARG(n == 0 || (x && dy), "NULL x/dy");
    AM_CHECK_ARG(workspace, "workspace is NULL");
    hipStream_t s = to_stream(stream);
    AM_MLP_DISPATCH(launch_bwd, a, x, n, params, dy, dx, dparams, workspace, workspace_bytes, s);
}
