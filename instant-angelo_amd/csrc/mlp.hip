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

// Output tiles of 16 rows the kernels run for n_out outputs (AM_MLP_DISPATCH
// picks the same): 1 up to 16 outputs, else 4.  The output layer's LDS rows
// are padded (with zeros) to all of them, since the backward's W^T * G reads
// every row of every tile.
static inline int mlp_out_tiles(int n_out) { return n_out > 16 ? 4 : 1; }

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
        r.n_rows_lds[L] = L == r.n_layers - 1 ? 16 * mlp_out_tiles(out) : 64;
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

// Forward of one 16-point tile (NH hidden layers); keeps the hidden
// pre-activations for the backward.
template <int KS, int OT, int NH>
__device__ __forceinline__ void tile_fwd(const MlpArgs& a, const float* sm, const float* xb,
                                         v4f (&z)[NH][4], v4f* zo) {
#pragma unroll
    for (int t = 0; t < 4; ++t) z[0][t] = mlp::bias_tile(sm + a.b_lds[0], t, 64);
    mlp::mm_fwd<KS, 4>(sm + a.w_lds[0], a.ld[0], mlp::FragF{xb}, z[0]);
#pragma unroll
    for (int L = 1; L < NH; ++L) {
        v4f h[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            h[t] = mlp::act4(a.act, z[L - 1][t], a.beta);
            z[L][t] = mlp::bias_tile(sm + a.b_lds[L], t, 64);
        }
        mlp::mm_fwd<16, 4>(sm + a.w_lds[L], a.ld[L], mlp::FragV{h}, z[L]);
    }
    v4f h[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) h[t] = mlp::act4(a.act, z[NH - 1][t], a.beta);
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

template <int KS, int OT, int NH>
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
        v4f z[NH][4], zo[OT];
        tile_fwd<KS, OT, NH>(a, sm, xb, z, zo);
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

// Writes a register dW tile set (OT x IT tiles; lane (i, k), register r ->
// row 16 ot + 4k + r, column 16 it + i) into a parameter-layout row.
template <int OT, int IT>
__device__ __forceinline__ void write_dw(float* out, int64_t w_par, int rows, int cols,
                                         const v4f* dW) {
    const int l = lane_id(), i = l & 15, k = l >> 4;
#pragma unroll
    for (int ot = 0; ot < OT; ++ot)
#pragma unroll
        for (int it = 0; it < IT; ++it)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int o = 16 * ot + 4 * k + r, c = 16 * it + i;
                if (o < rows && c < cols) out[w_par + (int64_t)o * cols + c] = dW[ot * IT + it][r];
            }
}

// Bias gradients: per-lane sums over the wave's samples, summed over the
// tile's 16 lanes at the end.
template <int T>
__device__ __forceinline__ void write_db(float* out, int64_t b_par, int rows, const v4f* g) {
    const int l = lane_id(), i = l & 15, k = l >> 4;
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const v4f sum = mlp::rowsum_tile(g[t]);
        if (i == 0)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                if (16 * t + 4 * k + r < rows) out[b_par + 16 * t + 4 * k + r] = sum[r];
    }
}

// Backward: waves of a block share the weights in LDS; each wave keeps its
// weight and bias gradients in registers over its tiles and writes them as
// one partial row (summed over the rows afterwards, a fixed order).
template <int KS, int OT, int NH>
__global__ void __launch_bounds__(256) k_mlp_bwd(MlpArgs a, const float* __restrict__ x, int64_t n,
                                                 const float* __restrict__ params,
                                                 const float* __restrict__ dy,
                                                 float* __restrict__ dx,
                                                 float* __restrict__ partial) {
    constexpr int IT0 = KS / 4;
    extern __shared__ float sm[];
    load_mlp_lds(a, params, sm);
    __syncthreads();
    const int l = lane_id(), i = l & 15, k = l >> 4;
    const int nw = blockDim.x >> 6, wave = threadIdx.x >> 6;
    float* scratch = sm + a.lds_floats + wave * (128 * mlp::kScratchLd);
    const int64_t n_tiles = (n + 15) / 16;
    const int Lo = NH;   // the output layer's index
    v4f dW0[4 * IT0], dWh[NH > 1 ? 16 * (NH - 1) : 1], dWo[4 * OT], db[NH][4], dbo[OT];
#pragma unroll
    for (int q = 0; q < 4 * IT0; ++q) dW0[q] = v4f{0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < (NH > 1 ? 16 * (NH - 1) : 1); ++q) dWh[q] = v4f{0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 4 * OT; ++q) dWo[q] = v4f{0, 0, 0, 0};
#pragma unroll
    for (int L = 0; L < NH; ++L)
#pragma unroll
        for (int t = 0; t < 4; ++t) db[L][t] = v4f{0, 0, 0, 0};
#pragma unroll
    for (int t = 0; t < OT; ++t) dbo[t] = v4f{0, 0, 0, 0};
    for (int64_t tile = (int64_t)blockIdx.x * nw + wave; tile < n_tiles;
         tile += (int64_t)gridDim.x * nw) {
        const int64_t p = tile * 16 + i;
        const bool valid = p < n;
        float xb[KS];
        load_x<KS>(a, x, p, valid, xb);
        v4f z[NH][4], zo[OT];
        tile_fwd<KS, OT, NH>(a, sm, xb, z, zo);
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
            for (int t = 0; t < 4; ++t) hin[t] = mlp::act4(a.act, z[NH - 1][t], a.beta);
            mlp::mm_dw<OT, 4, 16>(go, mlp::FragV{hin}, scratch, dWo);
#pragma unroll
            for (int t = 0; t < OT; ++t) dbo[t] += go[t];
#pragma unroll
            for (int t = 0; t < 4; ++t) gh[t] = v4f{0, 0, 0, 0};
            mlp::mm_bwd_in<4 * OT, 4>(sm + a.w_lds[Lo], a.ld[Lo], mlp::FragV{go}, gh);
        }
        // hidden layers, last to first
#pragma unroll
        for (int L = NH - 1; L >= 0; --L) {
            v4f gz[4];
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) gz[t][r] = gh[t][r] * mlp::act_bwd(a.act, z[L][t][r], a.beta);
#pragma unroll
            for (int t = 0; t < 4; ++t) db[L][t] += gz[t];
            if (L == 0) {
                mlp::mm_dw<4, IT0, KS>(gz, mlp::FragF{xb}, scratch, dW0);
                if (dx) {
                    v4f gx[IT0];
#pragma unroll
                    for (int t = 0; t < IT0; ++t) gx[t] = v4f{0, 0, 0, 0};
                    mlp::mm_bwd_in<16, IT0>(sm + a.w_lds[0], a.ld[0], mlp::FragV{gz}, gx);
                    if (valid)
#pragma unroll
                        for (int t = 0; t < IT0; ++t)
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
                mlp::mm_dw<4, 4, 16>(gz, mlp::FragV{hin}, scratch, dWh + 16 * (L - 1));
#pragma unroll
                for (int t = 0; t < 4; ++t) gh[t] = v4f{0, 0, 0, 0};
                mlp::mm_bwd_in<16, 4>(sm + a.w_lds[L], a.ld[L], mlp::FragV{gz}, gh);
            }
        }
    }
    // this wave's partial row in parameter layout (every wave writes all of it)
    float* out = partial + ((int64_t)blockIdx.x * nw + wave) * a.n_params;
    write_dw<4, IT0>(out, a.w_par[0], a.rows_out[0], a.cols_in[0], dW0);
    write_db<4>(out, a.b_par[0], a.rows_out[0], db[0]);
#pragma unroll
    for (int L = 1; L < NH; ++L) {
        write_dw<4, 4>(out, a.w_par[L], a.rows_out[L], a.cols_in[L], dWh + 16 * (L - 1));
        write_db<4>(out, a.b_par[L], a.rows_out[L], db[L]);
    }
    write_dw<OT, 4>(out, a.w_par[Lo], a.rows_out[Lo], a.cols_in[Lo], dWo);
    write_db<OT>(out, a.b_par[Lo], a.rows_out[Lo], dbo);
}

// Blocks of four waves; at most this many (the partial rows the column sum
// reads: waves x n_params floats).
static constexpr int kMlpBwdMaxBlocks = 512;

static size_t mlp_bwd_lds_bytes(const MlpArgs& a) {
    return (size_t)(a.lds_floats + 4 * 128 * mlp::kScratchLd) * sizeof(float);
}

static int mlp_bwd_waves(int64_t n) {
    return 4 * persistent_blocks((n + 15) / 16, 4, kMlpBwdMaxBlocks);
}

template <int KS, int OT, int NH>
static int launch_fwd(const MlpArgs& a, const float* x, int64_t n, const float* params, float* y,
                      hipStream_t s) {
    const size_t lds = (size_t)a.lds_floats * sizeof(float);
    AM_HIP(hipFuncSetAttribute((const void*)k_mlp_fwd<KS, OT, NH>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const int blocks = persistent_blocks((n + 15) / 16, 4, 2048);
    hipLaunchKernelGGL((k_mlp_fwd<KS, OT, NH>), dim3(blocks), dim3(256), lds, s, a, x, n, params,
                       y);
    AM_LAUNCH_CHECK();
    return AM_OK;
}

template <int KS, int OT, int NH>
static int launch_bwd(const MlpArgs& a, const float* x, int64_t n, const float* params,
                      const float* dy, float* dx, float* dparams, void* ws, size_t ws_bytes,
                      hipStream_t s) {
    const size_t lds = mlp_bwd_lds_bytes(a);
    AM_CHECK_ARG(lds <= 160 * 1024, "MLP too large for the LDS-resident backward (%zu bytes)", lds);
    const int waves = mlp_bwd_waves(n);
    WsCarver c(ws, ws_bytes);
    float* partial = c.take<float>((size_t)waves * a.n_params);
    float* tmp = c.take<float>((size_t)colsum_tmp_floats(waves, a.n_params) + 1);
    if (!c.ok()) return workspace_short(c.used, ws_bytes);
    AM_HIP(hipFuncSetAttribute((const void*)k_mlp_bwd<KS, OT, NH>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_mlp_bwd<KS, OT, NH>), dim3(waves / 4), dim3(256), lds, s, a, x, n, params,
                       dy, dx, partial);
    AM_LAUNCH_CHECK();
    return colsum_f32(partial, waves, a.n_params, dparams, tmp, s);
}

#define AM_MLP_DISPATCH_NH(FN, KS, OT, ...)                                  \
    do {                                                                     \
        switch (a.n_layers - 1) {                                            \
            case 1: return FN<KS, OT, 1>(__VA_ARGS__);                       \
            case 2: return FN<KS, OT, 2>(__VA_ARGS__);                       \
            case 3: return FN<KS, OT, 3>(__VA_ARGS__);                       \
            default: return FN<KS, OT, 4>(__VA_ARGS__);                      \
        }                                                                    \
    } while (0)

#define AM_MLP_DISPATCH(FN, ...)                                                          \
    do {                                                                                  \
        const int ks = 4 * (int)ceil_div(a.n_in, 16);                                     \
        const bool big = mlp_out_tiles(a.n_out) == 4;                                     \
        switch (ks) {                                                                     \
            case 4:                                                                       \
                if (big) AM_MLP_DISPATCH_NH(FN, 4, 4, __VA_ARGS__);                       \
                AM_MLP_DISPATCH_NH(FN, 4, 1, __VA_ARGS__);                                \
            case 8:                                                                       \
                if (big) AM_MLP_DISPATCH_NH(FN, 8, 4, __VA_ARGS__);                       \
                AM_MLP_DISPATCH_NH(FN, 8, 1, __VA_ARGS__);                                \
            case 12:                                                                      \
                if (big) AM_MLP_DISPATCH_NH(FN, 12, 4, __VA_ARGS__);                      \
                AM_MLP_DISPATCH_NH(FN, 12, 1, __VA_ARGS__);                               \
            default:                                                                      \
                if (big) AM_MLP_DISPATCH_NH(FN, 16, 4, __VA_ARGS__);                      \
                AM_MLP_DISPATCH_NH(FN, 16, 1, __VA_ARGS__);                               \
        }                                                                                 \
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
    const int waves = mlp_bwd_waves(n);
    WsCarver c(nullptr, 0);
    c.take<float>((size_t)waves * a.n_params);
    c.take<float>((size_t)colsum_tmp_floats(waves, a.n_params) + 1);
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
