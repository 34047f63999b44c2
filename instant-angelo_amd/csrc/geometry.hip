// Fused VolumeSDF.forward (models/geometry.py:195-287) on fp32 MFMA.
//
// One wave evaluates 16 samples.  Per sample: contraction (scale_anything,
// models/utils.py:109-114, with torch's CUDA scalar division = multiply by
// the fp32 reciprocal), the base evaluation, 6 central-difference
// evaluations (:219-234) and, with laplace, 6 around the tangent-shifted
// point (:236-275).  One evaluation = hash-grid encode (lane group k gathers
// levels k, k+4, k+8, k+12 of its sample, which are exactly its MFMA B
// fragments of the first layer) + the SDF MLP on MFMA.  The evaluations run
// in runtime loops so one copy of the encode + MLP body is compiled.
//
// The forward stores the 12 FD / curvature SDF values per sample and, in
// the store form, every evaluation's encoding.  The backward:
//   k_geo_bwd    rebuilds grad, normal, tangent and the curvature terms from
//                the stored values, re-runs each evaluation's MLP forward
//                (from the stored encoding, or a fresh gather in the
//                recompute form), back-propagates it on MFMA and writes the
//                encode gradient to HBM; MLP weight gradients are per-wave
//                register accumulators summed in a fixed order;
//   k_geo_pairs  turns the encode gradients into table-gradient pairs:
//     1. per lane and level, the evaluations of a cluster (base + 6 FD, or
//        the 6 curvature points) are merged into the cell B of the cluster
//        centre with a star stencil: an evaluation in B adds its 8 corners,
//        one in a face neighbour of B adds the 4 corners of the shared face
//        and emits its 4 far corners, any other emits its 8 corners;
//     2. B's corner sums are merged with equal-B lanes of the same lane
//        group (neighbouring samples of a ray) by a segmented reduction in a
//        fixed shuffle tree;
//     3. the same kernel in count mode makes these decisions from positions
//        alone first; an integer prefix sum of the per-tile counts gives
//        each tile its first pair, so the pairs land compacted, and a stable
//        radix sort + fixed-order segmented reduce gives dtable.
#include "hashgrid_dev.h"
#include "mlp_dev.h"
#include "table_grad.h"
#include "field_common.h"
#include <mutex>
#include <type_traits>
#include <rocprim/device/device_scan.hpp>

namespace am {
// Row stride (floats) of the 64-column weight matrices in LDS (W2 and the
// output layer): with 67, the row-major reads of the forward (lane i row i,
// lane group k column 4k) and the transposed reads of the backward (lane
// group k row 4k, lane i column i) both hit at most two lanes per bank; 65
// put four on one bank both ways.
constexpr int kWLd = 67;
}  // namespace am

namespace am {

using mlp::krow;

struct GeoArgs {
    GridArgs g;
    float r, inv2r, eps, inv_eps, inv_pi;
    int with_grad, with_laplace, recontract, active;
    int n_in, n_out, act;
    float beta;
    int w1, b1, w2, b2, wo, bo, ld1, lds_floats;  // LDS offsets (floats)
    // output tiles of 16 rows (n_out <= 16 * out_tiles); the rows of tiles
    // >= 1 are features only, read by the base evaluation alone
    int out_tiles;
    int64_t pw1, pb1, pw2, pb2, pwo, pbo, n_params;
    int n_evals;  // 1, 7 or 13
    int prec;     // AM_PREC_FP32 / AM_PREC_MIXED
    // analytic-gradient mode: the base evaluation carries d enc / d x and
    // grad = d sdf / d x through the MLP; no FD evaluations.  Slots: curv0 =
    // first curvature evaluation (7 FD mode, 1 analytic), n_dx = stored
    // position-derivative slots (the analytic base, then the 6 curvature
    // evaluations), genc_slots = encode-gradient slots per tile (n_evals, and
    // 3 more in analytic mode: u * scale * g_x along each axis).
    int analytic, curv0, n_dx, dx_curv0, genc_slots;
    // Evaluation order: kernel sample n is the caller's sample order[n]
    // (nullptr: the caller's order).  Inputs (positions, curvature
    // directions, upstream gradients) and outputs are addressed in the
    // caller's order; the forward's saved values (FD values, stored
    // encodings) and the encode gradients in the kernels' order.
    const int32_t* order;
};

__device__ __forceinline__ int64_t caller_row(const GeoArgs& a, int64_t n, bool valid) {
    return (a.order && valid) ? (int64_t)a.order[n] : n;
}

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
__device__ __forceinline__ void geo_load_lds(const GeoArgs& a, const float* __restrict__ params,
                                             float* sm, LevelParams* lv) {
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
            sm[a.w2 + o * kWLd + c] = params[a.pw2 + q];
        }
        for (int q = threadIdx.x; q < 64; q += blockDim.x) sm[a.b2 + q] = params[a.pb2 + q];
    }
    for (int q = threadIdx.x; q < a.out_tiles * 16 * 64; q += blockDim.x) {
        const int o = q >> 6, c = q & 63;
        sm[a.wo + o * kWLd + c] = o < a.n_out ? params[a.pwo + (int64_t)o * 64 + c] : 0.0f;
    }
    for (int q = threadIdx.x; q < a.out_tiles * 16; q += blockDim.x)
        sm[a.bo + q] = q < a.n_out ? params[a.pbo + q] : 0.0f;
    // unrolled: every kernel-argument access has a constant offset (a
    // per-thread index into the by-value struct would copy it to scratch)
#pragma unroll
    for (int q = 0; q < AM_MAX_LEVELS; ++q)
        if ((int)threadIdx.x == q && q < a.g.L) lv[q] = a.g.lv[q];
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

// Encode of this lane's level slots at contracted point x (e.xb[0..2LJ),
// and with DX the position derivatives e.dx).
template <int LJ, int NH, bool DX, class TT>
__device__ __forceinline__ void geo_encode(const GeoArgs& a, const LevelParams* lv,
                                           const TT* __restrict__ table, const float x[3],
                                           GeoEval<LJ, NH, DX>& e) {
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
}

// xyz input (last K-step) + SDF MLP forward from the encoded slots up to the
// last hidden pre-activation (all the backward needs).
template <int LJ, int NH, bool DX, int PR>
__device__ __forceinline__ void geo_mlp_hidden(const GeoArgs& a, const float* sm,
                                               const float x[3], GeoEval<LJ, NH, DX>& e) {
    constexpr int KS1 = GeoDims<LJ>::KS1;
    const int k = lane_id() >> 4;
    e.xb[2 * LJ] = k < 3 ? pick3(x, k) * 2.0f - 1.0f : 0.0f;
#pragma unroll
    for (int t = 0; t < 4; ++t) e.z1[t] = mlp::bias_tile(sm + a.b1, t, 64);
    mlp::mm_fwd<KS1, 4, mlp::FragF, PR>(sm + a.w1, a.ld1, mlp::FragF{e.xb}, e.z1);
    if constexpr (NH == 2) {
        v4f h[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) h[t] = mlp::act4(a.act, e.z1[t], a.beta);
#pragma unroll
        for (int t = 0; t < 4; ++t) e.z2[t] = mlp::bias_tile(sm + a.b2, t, 64);
        mlp::mm_fwd<16, 4, mlp::FragV, PR>(sm + a.w2, kWLd, mlp::FragV{h}, e.z2);
    }
}

// The full SDF MLP forward (hidden layers, then the output layer into e.out).
template <int LJ, int NH, bool DX, int PR>
__device__ __forceinline__ void geo_mlp(const GeoArgs& a, const float* sm, const float x[3],
                                        GeoEval<LJ, NH, DX>& e) {
    geo_mlp_hidden<LJ, NH, DX, PR>(a, sm, x, e);
    v4f h[4];
    const v4f* zl = NH == 2 ? e.z2 : e.z1;
#pragma unroll
    for (int t = 0; t < 4; ++t) h[t] = mlp::act4(a.act, zl[t], a.beta);
    e.out = mlp::bias_tile(sm + a.bo, 0, a.n_out);
    mlp::mm_fwd<16, 1, mlp::FragV, PR>(sm + a.wo, kWLd, mlp::FragV{h}, &e.out);
}

// Encode + SDF MLP of this lane's sample at contracted point x.
template <int LJ, int NH, bool DX, int PR, class TT>
__device__ __forceinline__ void geo_eval(const GeoArgs& a, const float* sm,
                                         const LevelParams* lv, const TT* __restrict__ table,
                                         const float x[3], GeoEval<LJ, NH, DX>& e) {
    geo_encode<LJ, NH, DX>(a, lv, table, x, e);
    geo_mlp<LJ, NH, DX, PR>(a, sm, x, e);
}

// Stored encodings (forward -> backward, the "store" form of the backward):
// per tile of 16 samples, evaluation ev (0 base, then 1-6 FD and 7-12
// curvature, or 1-6 curvature in analytic mode) and slot value s < 2LJ, one
// float per lane at
//   enc[((tile * n_evals + ev) * 2LJ + s) * 64 + lane];
// the position derivatives of the evaluations that carry them (dx slot c:
// the analytic base, then the curvature evaluations) follow at
// enc + geo_enc_floats(...):
//   [((tile * n_dx + c) * 6LJ + (j * 2 + f) * 3 + d) * 64 + lane].
// Every access is a 256 B coalesced row.
__host__ __device__ inline int64_t geo_enc_floats(int64_t tiles, int n_evals, int LJ) {
    return tiles * n_evals * 2 * LJ * 64;
}
__host__ __device__ inline int64_t geo_dx_floats(int64_t tiles, int n_dx, int LJ) {
    return tiles * n_dx * 6 * LJ * 64;
}

template <int LJ, int NH, bool DX>
__device__ __forceinline__ void geo_store(float* __restrict__ st, const GeoArgs& a, int64_t tile,
                                          int64_t n_tiles, int ev, int dxs,
                                          const GeoEval<LJ, NH, DX>& e) {
    const int l = lane_id();
    float* p = st + (tile * a.n_evals + ev) * (2 * LJ) * 64 + l;
#pragma unroll
    for (int q = 0; q < 2 * LJ; ++q) p[q * 64] = e.xb[q];
    if constexpr (DX) {
        float* d0 = st + geo_enc_floats(n_tiles, a.n_evals, LJ) +
                    (tile * a.n_dx + dxs) * (6 * LJ) * 64 + l;
#pragma unroll
        for (int j = 0; j < LJ; ++j)
#pragma unroll
            for (int f = 0; f < 2; ++f)
#pragma unroll
                for (int d = 0; d < 3; ++d) d0[((j * 2 + f) * 3 + d) * 64] = e.dx[j][f][d];
    }
}

template <int LJ, int NH, bool DX>
__device__ __forceinline__ void geo_load(const float* __restrict__ st, const GeoArgs& a,
                                         int64_t tile, int64_t n_tiles, int ev, int dxs,
                                         GeoEval<LJ, NH, DX>& e) {
    const int l = lane_id();
    const float* p = st + (tile * a.n_evals + ev) * (2 * LJ) * 64 + l;
#pragma unroll
    for (int q = 0; q < 2 * LJ; ++q) e.xb[q] = p[q * 64];
    if constexpr (DX) {
        const float* d0 = st + geo_enc_floats(n_tiles, a.n_evals, LJ) +
                          (tile * a.n_dx + dxs) * (6 * LJ) * 64 + l;
#pragma unroll
        for (int j = 0; j < LJ; ++j)
#pragma unroll
            for (int f = 0; f < 2; ++f)
#pragma unroll
                for (int d = 0; d < 3; ++d) e.dx[j][f][d] = d0[((j * 2 + f) * 3 + d) * 64];
    }
}

// Backward-side evaluation: the stored encoding (ST) or a fresh encode, then
// the MLP forward up to the last hidden pre-activation.
template <int LJ, int NH, bool DX, bool ST, int PR, class TT>
__device__ __forceinline__ void geo_eval_b(const GeoArgs& a, const float* sm, const LevelParams* lv,
                                           const TT* __restrict__ table,
                                           const float* __restrict__ st, int64_t tile,
                                           int64_t n_tiles, int ev, int dxs, const float x[3],
                                           GeoEval<LJ, NH, DX>& e) {
    if constexpr (ST) geo_load<LJ, NH, DX>(st, a, tile, n_tiles, ev, dxs, e);
    else geo_encode<LJ, NH, DX>(a, lv, table, x, e);
    geo_mlp_hidden<LJ, NH, DX, PR>(a, sm, x, e);
}

__device__ __forceinline__ float contract1(float v, const GeoArgs& a) { return (v + a.r) * a.inv2r; }

__device__ __forceinline__ float clampr(float v, float r) { return fminf(fmaxf(v, -r), r); }

// Offset point j (axis j/2, sign + for even j) around `base` (world units,
// as the reference adds the offsets to it), clamped and contracted; mask =
// the clamp passed the coordinate through.
__device__ __forceinline__ void offset_point(const GeoArgs& a, const float base[3], int j,
                                             float x[3], bool mask[3]) {
    const int ax = j >> 1;
    const float off = (j & 1) ? -a.eps : a.eps;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float v = base[d] + (d == ax ? off : 0.0f);
        mask[d] = v >= -a.r && v <= a.r;
        x[d] = contract1(clampr(v, a.r), a);
    }
}

__device__ __forceinline__ void normalize3(const float v[3], float out[3]) {
    const float n = sqrtf((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    const float d = fmaxf(n, 1e-12f);
#pragma unroll
    for (int q = 0; q < 3; ++q) out[q] = v[q] / d;
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

// Per-sample state rebuilt from the 12 stored SDF values (all lanes of the
// sample hold it).
struct GeoSample {
    float P[3], x0[3];
    float grad[3], n[3];
    float rn[3], ps[3], grad_s[3], ns[3], dot;
};

__device__ __forceinline__ void diff_pair(float* acc3, int j, float v) {
    // acc[j/2] = f+ then f+ - f- : bit-identical to the single subtraction
    const int ax = j >> 1;
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if (d == ax) acc3[d] = (j & 1) ? acc3[d] - v : v;
}

__device__ __forceinline__ void finish_grad(const GeoArgs& a, float g[3]) {
#pragma unroll
    for (int d = 0; d < 3; ++d) g[d] = (0.5f * g[d]) * a.inv_eps;
}

// grad -> normal -> tangent -> shifted point.
__device__ __forceinline__ void geo_tangent(const GeoArgs& a, const float* __restrict__ randn,
                                            int64_t n, bool valid, GeoSample& s) {
    float rd[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) rd[d] = valid ? randn[3 * n + d] : 1.0f;
    normalize3(rd, s.rn);
    float t[3];
    cross3(s.n, s.rn, t);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float base = a.recontract ? s.x0[d] : s.P[d];
        s.ps[d] = base + t[d] * a.eps;
    }
}

// Precision forms of the fused field (template PR).  Mixed: the table rows
// are read in fp16 and the table-gradient pair values are stored in fp16
// (summed in fp32); the SDF MLP stays fp32 in both forms, because the
// central differences of the FD normal divide each evaluation's rounding
// by eps (bf16 activations there gave normals off by 0.3 on average at the
// Neuralangelo shapes, while a fixed fp16 rounding of the table keeps the
// field smooth).
// Table storage of a precision form: fp32 rows, or fp16 rows (mixed form).
template <int PR>
using TableT = typename std::conditional<PR == mlp::kF32, float2, __half2>::type;
// Table-gradient pair values: fp32, or fp16 (mixed form, loss-scaled).
template <int PR>
using PairV = typename std::conditional<PR == mlp::kF32, Vals<2>, ValsH<2>>::type;

// sdf and feature outputs of the base evaluation: output tile 0 from e.out,
// the feature rows of tiles >= 1 (n_out > 16) from the last hidden layer.
template <int LJ, int NH, bool DX>
__device__ __forceinline__ void write_base(const GeoArgs& a, const float* sm,
                                           const GeoEval<LJ, NH, DX>& e, const GeoSample& s,
                                           int64_t n, bool valid, int k, int fdim,
                                           float* __restrict__ sdf, float* __restrict__ feature) {
    // (wide outputs come with two hidden layers: make_geo_args rejects them
    // with one, whose kernels keep their registers for the narrow form)
    if (NH == 2 && feature && a.out_tiles > 1) {   // every lane takes part in the MFMA
        v4f h[4];
        const v4f* zl = NH == 2 ? e.z2 : e.z1;
#pragma unroll
        for (int t = 0; t < 4; ++t) h[t] = mlp::act4(a.act, zl[t], a.beta);
#pragma unroll 1
        for (int t = 1; t < a.out_tiles; ++t) {
            v4f o = mlp::bias_tile(sm + a.bo, t, a.n_out);
            mlp::mm_fwd<16, 1, mlp::FragV, mlp::kF32>(sm + a.wo + 16 * t * kWLd, kWLd, mlp::FragV{h},
                                                      &o);
            if (valid)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * t + 4 * k + r;
                    if (row < a.n_out) feature[n * fdim + row] = o[r];
                }
        }
    }
    if (!valid) return;
    if (k == 0) sdf[n] = e.out[0];
    if (feature) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 4 * k + r;
            if (row < a.n_out) feature[n * fdim + row] = e.out[r];
        }
        if (k < 3) feature[n * fdim + a.n_out + k] = pick3(s.x0, k) * 2.0f - 1.0f;
    }
}

// Row 0 of the output layer (the SDF row), Wo[0, hidden], in D-tile layout.
__device__ __forceinline__ void sdf_row_tiles(const GeoArgs& a, const float* sm, v4f w0[4]) {
    const int k = lane_id() >> 4;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) w0[t][r] = sm[a.wo + 16 * t + 4 * k + r];
}

// d sdf / d x01 of the base evaluation (analytic mode): u = W1^T c1, with
// c1 = sigma'(z1) * Wo[0,:] (one hidden layer) or c1 = sigma'(z1) * W2^T
// (sigma'(z2) * Wo[0,:]) (two), is the gradient of the SDF row with respect
// to the MLP input (encoding slots, then xyz*2-1), contracted with the
// encoding's position derivatives e.dx and with 2 for the xyz input; summed
// over the four lane groups (every lane of the sample gets the sum).  u is
// returned in B-fragment layout.
template <int LJ, int NH>
__device__ __forceinline__ void geo_input_grad(const GeoArgs& a, const float* sm,
                                               const GeoEval<LJ, NH, true>& e, float gpos[3],
                                               float u[GeoDims<LJ>::KS1]) {
    constexpr int IT1 = GeoDims<LJ>::IT1;
    const int k = lane_id() >> 4;
    v4f c1[4];
    sdf_row_tiles(a, sm, c1);
    if constexpr (NH == 2) {
        v4f c2[4], e1[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            v4f h, dh;
            mlp::act_fb4(a.act, e.z2[t], a.beta, h, dh);
            c2[t] = dh * c1[t];
            e1[t] = v4f{0, 0, 0, 0};
        }
        mlp::mm_bwd_in<16, 4, mlp::FragV, mlp::kF32>(sm + a.w2, kWLd, mlp::FragV{c2}, e1);
#pragma unroll
        for (int t = 0; t < 4; ++t) c1[t] = e1[t];
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        v4f h, dh;
        mlp::act_fb4(a.act, e.z1[t], a.beta, h, dh);
        c1[t] = dh * c1[t];
    }
    v4f gin[IT1];
#pragma unroll
    for (int t = 0; t < IT1; ++t) gin[t] = v4f{0, 0, 0, 0};
    mlp::mm_bwd_in<16, IT1, mlp::FragV, mlp::kF32>(sm + a.w1, a.ld1, mlp::FragV{c1}, gin);
    const mlp::FragV gb{gin};
#pragma unroll
    for (int q = 0; q < GeoDims<LJ>::KS1; ++q) u[q] = gb[q];
    float gp[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int j = 0; j < LJ; ++j)
#pragma unroll
        for (int d = 0; d < 3; ++d) gp[d] += gb[2 * j] * e.dx[j][0][d] + gb[2 * j + 1] * e.dx[j][1][d];
    if (k < 3) {
        const float gxyz = gb[2 * LJ] * 2.0f;
#pragma unroll
        for (int d = 0; d < 3; ++d) gp[d] += d == k ? gxyz : 0.0f;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) gpos[d] = sum_groups(gp[d]);
}

// Forward of the fused field.  AN: analytic-gradient mode (grad from the
// base evaluation's derivatives; one hidden layer), else central
// differences.  fdv keeps, per sample, the 6 FD values (or the analytic
// grad in [0, 3)) and the 6 curvature values, for the backward.
template <int LJ, int NH, bool ST, int PR, bool AN>
__global__ void __launch_bounds__(256) k_geo_fwd(GeoArgs a, const float* __restrict__ pts,
                                                 const float* __restrict__ randn, int64_t N,
                                                 const void* __restrict__ table_v,
                                                 const float* __restrict__ params,
                                                 float* __restrict__ sdf, float* __restrict__ grad,
                                                 float* __restrict__ normal,
                                                 float* __restrict__ feature,
                                                 float* __restrict__ laplace,
                                                 float* __restrict__ fdv,
                                                 float* __restrict__ store) {
    extern __shared__ float sm[];
    __shared__ LevelParams lv[AM_MAX_LEVELS];
    geo_load_lds<LJ, NH>(a, params, sm, lv);
    __syncthreads();
    const TableT<PR>* table = reinterpret_cast<const TableT<PR>*>(table_v);
    const int l = lane_id(), p = l & 15, k = l >> 4;
    const int nw = blockDim.x >> 6, wave = threadIdx.x >> 6;
    const int64_t n_tiles = (N + 15) / 16;
    const int fdim = a.n_out + 3;
    const int vb = xcd_block((int)blockIdx.x, (int)gridDim.x);   // the table gathers share L2
    for (int64_t tile = (int64_t)vb * nw + wave; tile < n_tiles;
         tile += (int64_t)gridDim.x * nw) {
        const int64_t n = tile * 16 + p;
        const bool valid = n < N;
        const int64_t m = caller_row(a, n, valid);
        GeoSample s;
#pragma unroll
        for (int d = 0; d < 3; ++d) s.P[d] = valid ? pts[3 * m + d] : 0.0f;
#pragma unroll
        for (int d = 0; d < 3; ++d) s.x0[d] = contract1(s.P[d], a);
        if constexpr (AN) {
            GeoEval<LJ, NH, true> e;   // with the position derivatives
            geo_eval<LJ, NH, true, mlp::kF32>(a, sm, lv, table, s.x0, e);
            if constexpr (ST) geo_store<LJ, NH, true>(store, a, tile, n_tiles, 0, 0, e);
            write_base(a, sm, e, s, m, valid, k, fdim, sdf, feature);
            if (!a.with_grad) continue;
            float gp[3], u[GeoDims<LJ>::KS1];
            geo_input_grad<LJ, NH>(a, sm, e, gp, u);
#pragma unroll
            for (int d = 0; d < 3; ++d) s.grad[d] = gp[d] * a.inv2r;   // d x01 / d x = 1 / 2r
            if (fdv && valid && k == 0)
#pragma unroll
                for (int d = 0; d < 3; ++d) fdv[n * 12 + d] = s.grad[d];
        } else {
            {
                GeoEval<LJ, NH, false> e;
                geo_eval<LJ, NH, false, mlp::kF32>(a, sm, lv, table, s.x0, e);
                if constexpr (ST) geo_store<LJ, NH, false>(store, a, tile, n_tiles, 0, 0, e);
                write_base(a, sm, e, s, m, valid, k, fdim, sdf, feature);
            }
            if (!a.with_grad) continue;
            float g[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll 1
            for (int j = 0; j < 6; ++j) {
                float x[3];
                bool inb[3];
                offset_point(a, s.P, j, x, inb);
                GeoEval<LJ, NH, false> e;
                geo_eval<LJ, NH, false, mlp::kF32>(a, sm, lv, table, x, e);
                if constexpr (ST) geo_store<LJ, NH, false>(store, a, tile, n_tiles, 1 + j, 0, e);
                const float v = shfl(e.out[0], p);
                diff_pair(g, j, v);
                if (fdv && valid && k == 0) fdv[n * 12 + j] = v;
            }
            finish_grad(a, g);
#pragma unroll
            for (int d = 0; d < 3; ++d) s.grad[d] = g[d];
        }
        normalize3(s.grad, s.n);
        if (valid && k == 0) {
            if (grad)
#pragma unroll
                for (int d = 0; d < 3; ++d) grad[3 * m + d] = s.grad[d];
            if (normal)
#pragma unroll
                for (int d = 0; d < 3; ++d) normal[3 * m + d] = s.n[d];
        }
        if (!a.with_laplace) continue;
        geo_tangent(a, randn, m, valid, s);
        float gs[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll 1
        for (int j = 0; j < 6; ++j) {
            float x[3];
            bool inb[3];
            offset_point(a, s.ps, j, x, inb);
            GeoEval<LJ, NH, ST> e;  // the store keeps the position derivatives
            geo_eval<LJ, NH, ST, mlp::kF32>(a, sm, lv, table, x, e);
            if constexpr (ST)
                geo_store<LJ, NH, true>(store, a, tile, n_tiles, a.curv0 + j, a.dx_curv0 + j, e);
            const float v = shfl(e.out[0], p);
            diff_pair(gs, j, v);
            if (fdv && valid && k == 0) fdv[n * 12 + 6 + j] = v;
        }
        finish_grad(a, gs);
        float ns[3];
        normalize3(gs, ns);
        const float dot = (s.n[0] * ns[0] + s.n[1] * ns[1]) + s.n[2] * ns[2];
        if (valid && k == 0 && laplace) {
            const float lo = -1.0f + 1e-6f, hi = 1.0f - 1e-6f;
            laplace[m] = acosf(fminf(fmaxf(dot, lo), hi)) * a.inv_pi;
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

// Feature rows of output tiles >= 1 (n_out > 16) in the base evaluation's
// backward: their upstream gradients come from g_feature, their weight and
// bias gradients accumulate in a wave-private LDS area (acc) instead of
// registers: per tile t >= 1, [4 column tiles][64 lanes] v4f of dWo (the
// mm_dw layout) then 16 row sums of dbo.
struct OutExt {
    const float* g_feature;
    int64_t m;   // the sample's caller row
    int fdim;
    bool valid;
    float* acc;
};
constexpr int kOutExtFloats = 4 * 64 * 4 + 16;   // per extra tile

__device__ __forceinline__ void out_ext_bwd(const GeoArgs& a, const float* sm, const v4f h[4],
                                            const OutExt& x, float* scratch, v4f gh[4]) {
    const int l = lane_id(), i = l & 15, k = l >> 4;
#pragma unroll 1
    for (int t = 1; t < a.out_tiles; ++t) {
        v4f got;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * t + 4 * k + r;
            got[r] = (x.valid && x.g_feature && row < a.n_out) ? x.g_feature[x.m * x.fdim + row]
                                                                 : 0.0f;
        }
        float* acc = x.acc + (t - 1) * kOutExtFloats;
        v4f dw[4];
#pragma unroll
        for (int it = 0; it < 4; ++it) dw[it] = *reinterpret_cast<const v4f*>(acc + (it * 64 + l) * 4);
        mlp::mm_dw<1, 4, 16, mlp::FragV, mlp::kF32>(&got, mlp::FragV{h}, scratch, dw);
#pragma unroll
        for (int it = 0; it < 4; ++it) *reinterpret_cast<v4f*>(acc + (it * 64 + l) * 4) = dw[it];
        const v4f bs = mlp::rowsum_tile(got);
        if (i == 0)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[1024 + 4 * k + r] += bs[r];
        mlp::mm_bwd_in<4, 4, mlp::FragV, mlp::kF32>(sm + a.wo + 16 * t * kWLd, kWLd, mlp::FragV{&got},
                                                    gh);
    }
}

// LDS float offset of the backward's per-wave OutExt areas (after the
// weights and the four waves' transpose scratch; 16-byte aligned).
__host__ __device__ inline int geo_ext_base(const GeoArgs& a) {
    return (a.lds_floats + 4 * 128 * mlp::kScratchLd + 3) & ~3;
}

// Output-gradient tile with only row 0 (the SDF) set.
__device__ __forceinline__ v4f go_row0(float g) {
    const int k = lane_id() >> 4;
    return v4f{k == 0 ? g : 0.0f, 0.0f, 0.0f, 0.0f};
}

// MLP backward of one evaluation with output gradient `go` (D tile rows =
// outputs): accumulates weight gradients, stores the encode gradient of the
// lane's level slots to genc (LDS, slot-major [LJ][2][64]) and, with DX,
// returns the contracted-position gradient (summed over the 4 lane groups).
template <int LJ, int NH, bool DX, int PR>
__device__ __forceinline__ void geo_eval_bwd(const GeoArgs& a, const float* sm,
                                             const GeoEval<LJ, NH, DX>& e, v4f go,
                                             float* scratch, GeoGrad<LJ, NH>& G, float* genc,
                                             float gpos[3], const OutExt* ext = nullptr) {
    constexpr int KS1 = GeoDims<LJ>::KS1, IT1 = GeoDims<LJ>::IT1;
    const int l = lane_id(), k = l >> 4;
    v4f h[4], dh[4];   // last hidden activation and its derivative
    const v4f* zl = NH == 2 ? e.z2 : e.z1;
#pragma unroll
    for (int t = 0; t < 4; ++t) mlp::act_fb4(a.act, zl[t], a.beta, h[t], dh[t]);
    mlp::mm_dw<1, 4, 16, mlp::FragV, PR>(&go, mlp::FragV{h}, scratch, G.dWo);
    G.dbo += go;
    v4f gh[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) gh[t] = v4f{0, 0, 0, 0};
    mlp::mm_bwd_in<4, 4, mlp::FragV, PR>(sm + a.wo, kWLd, mlp::FragV{&go}, gh);
    if constexpr (NH == 2)
        if (ext) out_ext_bwd(a, sm, h, *ext, scratch, gh);
    if constexpr (NH == 2) {
        v4f gz2[4], h1[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            gz2[t] = gh[t] * dh[t];
            mlp::act_fb4(a.act, e.z1[t], a.beta, h1[t], dh[t]);
            G.db2[t] += gz2[t];
        }
        mlp::mm_dw<4, 4, 16, mlp::FragV, PR>(gz2, mlp::FragV{h1}, scratch, G.dW2);
#pragma unroll
        for (int t = 0; t < 4; ++t) gh[t] = v4f{0, 0, 0, 0};
        mlp::mm_bwd_in<16, 4, mlp::FragV, PR>(sm + a.w2, kWLd, mlp::FragV{gz2}, gh);
    }
    v4f gz1[4];   // dh now holds the first hidden layer's derivative
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        gz1[t] = gh[t] * dh[t];
        G.db1[t] += gz1[t];
    }
    mlp::mm_dw<4, IT1, KS1, mlp::FragF, PR>(gz1, mlp::FragF{e.xb}, scratch, G.dW1);
    v4f gin[IT1];
#pragma unroll
    for (int t = 0; t < IT1; ++t) gin[t] = v4f{0, 0, 0, 0};
    mlp::mm_bwd_in<16, IT1, mlp::FragV, PR>(sm + a.w1, a.ld1, mlp::FragV{gz1}, gin);
    const mlp::FragV gb{gin};  // gb[s] = gradient of B fragment s
#pragma unroll
    for (int j = 0; j < LJ; ++j) {
        genc[(2 * j) * 64 + l] = gb[2 * j];
        genc[(2 * j + 1) * 64 + l] = gb[2 * j + 1];
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

// Backward of the analytic base evaluation.  With v = d input / d x01 . g_x
// (the encoding's position derivatives and 2 for xyz, contracted with the
// upstream gradient of d sdf / d x01), the gradient term of the loss is the
// directional derivative of the SDF row along v: a tangent pass
// s1 = W1 v, a1' = sigma'(z1) s1 (, s2 = W2 a1', a2' = sigma'(z2) s2) and
// sdf' = Wo[0,:] . a_last'.  Its reverse pass, with the ordinary output
// gradient go alongside (adjoints zb of the pre-activations and sb of the
// tangents):
//   dWo += go (x) a_last + e0 (x) a_last'      (e0: row 0 = 1)
//   layer l from the top: sb_l = sigma'(z_l) * ab'_l,
//     zb_l = sigma'(z_l) * ab_l + sigma''(z_l) * s_l * ab'_l,
//     dW_l += zb_l (x) a_{l-1} + sb_l (x) a'_{l-1},  db_l += zb_l,
//     ab_{l-1} = W_l^T zb_l,  ab'_{l-1} = W_l^T sb_l
//   (ab_last = Wo^T go, ab'_last = Wo[0,:]; a_0 = the input, a'_0 = v).
// The input adjoint W1^T zb_1 is the ordinary encode gradient (genc); the
// tangent adjoint u = W1^T sb_1 reaches the table through v: each corner c
// of level l gets u[l,f] * scale_l * grad(w_c) . g_x, stored as
// u[l,f] * scale_l * g_x[d] in three extra slots (gext) for the pair pass.
template <int LJ, int NH>
__device__ __forceinline__ void geo_eval_bwd_an(const GeoArgs& a, const float* sm,
                                                const LevelParams* lv,
                                                const GeoEval<LJ, NH, true>& e, v4f go,
                                                const float gx[3], float* scratch,
                                                GeoGrad<LJ, NH>& G, float* genc, float* gext,
                                                const OutExt* ext) {
    constexpr int KS1 = GeoDims<LJ>::KS1, IT1 = GeoDims<LJ>::IT1;
    const int l = lane_id(), k = l >> 4;
    float v[KS1];
#pragma unroll
    for (int j = 0; j < LJ; ++j)
#pragma unroll
        for (int f = 0; f < 2; ++f)
            v[2 * j + f] = (e.dx[j][f][0] * gx[0] + e.dx[j][f][1] * gx[1]) + e.dx[j][f][2] * gx[2];
    v[2 * LJ] = k < 3 ? 2.0f * pick3(gx, k) : 0.0f;
    v4f h1[4], dh1[4], sv1[4], w0[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        mlp::act_fb4(a.act, e.z1[t], a.beta, h1[t], dh1[t]);
        sv1[t] = v4f{0, 0, 0, 0};
    }
    sdf_row_tiles(a, sm, w0);
    mlp::mm_fwd<KS1, 4, mlp::FragF, mlp::kF32>(sm + a.w1, a.ld1, mlp::FragF{v}, sv1);
    v4f gz1[4], c1[4];   // zb_1 and sb_1
    if constexpr (NH == 1) {
        // output layer: the ordinary term and the gradient term's row 0
        mlp::mm_dw<1, 4, 16, mlp::FragV, mlp::kF32>(&go, mlp::FragV{h1}, scratch, G.dWo);
        G.dbo += go;
        {
            v4f sdh[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) sdh[t] = sv1[t] * dh1[t];
            const v4f e0 = go_row0(1.0f);
            mlp::mm_dw<1, 4, 16, mlp::FragV, mlp::kF32>(&e0, mlp::FragV{sdh}, scratch, G.dWo);
        }
        v4f gh[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) gh[t] = v4f{0, 0, 0, 0};
        mlp::mm_bwd_in<4, 4, mlp::FragV, mlp::kF32>(sm + a.wo, kWLd, mlp::FragV{&go}, gh);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            v4f d2;
#pragma unroll
            for (int r = 0; r < 4; ++r) d2[r] = mlp::act_d2(a.act, e.z1[t][r], a.beta, dh1[t][r]);
            gz1[t] = gh[t] * dh1[t] + sv1[t] * w0[t] * d2;
            c1[t] = dh1[t] * w0[t];
            G.db1[t] += gz1[t];
        }
    } else {
        // tangent of the second layer
        v4f ad1[4], sv2[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            ad1[t] = dh1[t] * sv1[t];
            sv2[t] = v4f{0, 0, 0, 0};
        }
        mlp::mm_fwd<16, 4, mlp::FragV, mlp::kF32>(sm + a.w2, kWLd, mlp::FragV{ad1}, sv2);
        v4f gz2[4], c2[4];
        {
            v4f h2[4], dh2[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) mlp::act_fb4(a.act, e.z2[t], a.beta, h2[t], dh2[t]);
            mlp::mm_dw<1, 4, 16, mlp::FragV, mlp::kF32>(&go, mlp::FragV{h2}, scratch, G.dWo);
            G.dbo += go;
            {
                v4f ad2[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) ad2[t] = dh2[t] * sv2[t];
                const v4f e0 = go_row0(1.0f);
                mlp::mm_dw<1, 4, 16, mlp::FragV, mlp::kF32>(&e0, mlp::FragV{ad2}, scratch, G.dWo);
            }
            v4f gh[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) gh[t] = v4f{0, 0, 0, 0};
            mlp::mm_bwd_in<4, 4, mlp::FragV, mlp::kF32>(sm + a.wo, kWLd, mlp::FragV{&go}, gh);
            if (ext) out_ext_bwd(a, sm, h2, *ext, scratch, gh);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                v4f d2;
#pragma unroll
                for (int r = 0; r < 4; ++r) d2[r] = mlp::act_d2(a.act, e.z2[t][r], a.beta, dh2[t][r]);
                gz2[t] = gh[t] * dh2[t] + sv2[t] * w0[t] * d2;
                c2[t] = dh2[t] * w0[t];
                G.db2[t] += gz2[t];
            }
        }
        mlp::mm_dw<4, 4, 16, mlp::FragV, mlp::kF32>(gz2, mlp::FragV{h1}, scratch, G.dW2);
        mlp::mm_dw<4, 4, 16, mlp::FragV, mlp::kF32>(c2, mlp::FragV{ad1}, scratch, G.dW2);
        v4f gh1[4], e1[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            gh1[t] = v4f{0, 0, 0, 0};
            e1[t] = v4f{0, 0, 0, 0};
        }
        mlp::mm_bwd_in<16, 4, mlp::FragV, mlp::kF32>(sm + a.w2, kWLd, mlp::FragV{gz2}, gh1);
        mlp::mm_bwd_in<16, 4, mlp::FragV, mlp::kF32>(sm + a.w2, kWLd, mlp::FragV{c2}, e1);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            v4f d2;
#pragma unroll
            for (int r = 0; r < 4; ++r) d2[r] = mlp::act_d2(a.act, e.z1[t][r], a.beta, dh1[t][r]);
            gz1[t] = gh1[t] * dh1[t] + sv1[t] * e1[t] * d2;
            c1[t] = dh1[t] * e1[t];
            G.db1[t] += gz1[t];
        }
    }
    mlp::mm_dw<4, IT1, KS1, mlp::FragF, mlp::kF32>(gz1, mlp::FragF{e.xb}, scratch, G.dW1);
    mlp::mm_dw<4, IT1, KS1, mlp::FragF, mlp::kF32>(c1, mlp::FragF{v}, scratch, G.dW1);
    v4f gin[IT1], uin[IT1];
#pragma unroll
    for (int t = 0; t < IT1; ++t) {
        gin[t] = v4f{0, 0, 0, 0};
        uin[t] = v4f{0, 0, 0, 0};
    }
    mlp::mm_bwd_in<16, IT1, mlp::FragV, mlp::kF32>(sm + a.w1, a.ld1, mlp::FragV{gz1}, gin);
    mlp::mm_bwd_in<16, IT1, mlp::FragV, mlp::kF32>(sm + a.w1, a.ld1, mlp::FragV{c1}, uin);
    const mlp::FragV gb{gin}, ub{uin};
#pragma unroll
    for (int j = 0; j < LJ; ++j) {
        genc[(2 * j) * 64 + l] = gb[2 * j];
        genc[(2 * j + 1) * 64 + l] = gb[2 * j + 1];
        const int level = k + 4 * j;
        const float sc = level < a.g.L ? lv[level].scale : 0.0f;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            gext[d * LJ * 128 + (2 * j) * 64 + l] = ub[2 * j] * sc * gx[d];
            gext[d * LJ * 128 + (2 * j + 1) * 64 + l] = ub[2 * j + 1] * sc * gx[d];
        }
    }
}

// ---- pair emission with merging ----
// A tile's pairs go to [keys, keys + limit): limit is the count the count
// pass found for the tile; the emit pass makes the same decisions, and never
// writes past limit.
template <class VT>
struct EmitTile {
    uint32_t* keys;   // tile's first pair (keys); partitioned: the arrays' start
    VT* vals;         // tile's first pair (values); partitioned: the arrays' start
    int cursor;       // wave-uniform write position
    int limit;        // pairs reserved for the tile
    // Partitioned form (part != nullptr): the pairs of the tile with radix
    // digit d = (row >> 9) & 255 go to base[d] + [0, lim[d]); part[d] counts
    // them as they are written (in the count pass: the counts themselves).
    // Wave-private LDS arrays of kPartDigits entries.
    int32_t* part;
    const int32_t* lim;
    const int64_t* base;
};

// Partitioned pair emission (flag "table_prepartition", off by default): the
// first 8-bit digit of the table-gradient radix sort (key bits [9, 17)) is
// applied as pairs are written, and the sort makes one pass over [17, 25)
// instead of two.  Measured on the Neuralangelo step it loses: the sort
// drops from 4.6 to 2.4 ms, but the emission's stores scatter over 256
// partitions and write partial lines (emit 3.0 -> 7.7 ms, count 0.4 ->
// 0.9 ms); 32 partitions with one 11-bit pass measured no better (emit 5.5,
// sort 4.7 ms).
constexpr int kPartShift = 9;
constexpr int kPartBits = 8;
constexpr int kPartDigits = 1 << kPartBits;

__device__ __forceinline__ uint32_t pair_digit(uint32_t row) {
    return (row >> kPartShift) & (uint32_t)(kPartDigits - 1);
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int l = lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Atomic form (flag "table_atomic"): the "pair array" is the table gradient
// itself, [n_entries] rows of 2 floats, and each pair is added into its row
// with hardware float atomics: no count pass, no sort, no host read, but the
// summation order varies from run to run (the tcnn hash-grid backward's
// semantics).
struct AtomicRow {
    float v[2];
};
template <class VT>
constexpr bool kAtomicPairs = std::is_same<VT, AtomicRow>::value;

// Lanes with `on` write K pairs (rows, values); COUNT only advances the
// cursor.
__device__ __forceinline__ void set_pair(Vals<2>& v, float a, float b) {
    v.v[0] = a;
    v.v[1] = b;
}
__device__ __forceinline__ void set_pair(ValsH<2>& v, float a, float b) {
    v.v[0] = __float2half_rn(a);
    v.v[1] = __float2half_rn(b);
}
__device__ __forceinline__ void set_pair(AtomicRow& v, float a, float b) {   // type-checks only
    v.v[0] = a;
    v.v[1] = b;
}

// The group's K * n pairs (n lanes on) fill [cursor, cursor + K n) corner-
// major: pair c of the lane of rank r goes to cursor + c n + r, so each store
// instruction writes consecutive addresses across the lanes.  The order is a
// fixed function of the data (the reduction's summation order follows it).
template <int K, bool COUNT, class VT>
__device__ __forceinline__ void emit_pairs(EmitTile<VT>& et, bool on, const uint32_t rows[K],
                                           const float vals[2 * K]) {
    if constexpr (kAtomicPairs<VT>) {
        static_assert(!COUNT, "the atomic form has no count pass");
        if (on) {
#pragma unroll
            for (int c = 0; c < K; ++c) {
                float* row = et.vals[rows[c]].v;
                unsafeAtomicAdd(row, vals[2 * c]);
                unsafeAtomicAdd(row + 1, vals[2 * c + 1]);
            }
        }
        return;
    }
    const uint64_t mask = __ballot(on);
    const int n_on = __popcll(mask);
    if (et.part) {   // partitioned form (rows are computed in both passes)
        if (mask == 0) return;
        const uint64_t lt = lanemask_lt();
#pragma unroll
        for (int c = 0; c < K; ++c) {
            const uint32_t d = on ? pair_digit(rows[c]) : 0u;
            if constexpr (COUNT) {
                if (on) atomicAdd(&et.part[d], 1);   // counts: order-free
            } else {
                // lanes with the same digit (one ballot a bit), ranked by lane: the
                // order inside a partition is corner-major, then lane, as in
                // the compact form, so every sum keeps its order
                uint64_t peers = mask;
#pragma unroll
                for (int b = 0; b < kPartBits; ++b) {
                    const uint64_t ones = __ballot(on && ((d >> b) & 1u));
                    peers &= ((d >> b) & 1u) ? ones : ~ones;
                }
                const int before = __popcll(peers & lt);
                // one wave's LDS operations complete in issue order: every
                // lane reads part[d] before its group's first lane updates it
                const int have = on ? et.part[d] : 0;
                if (on && before == 0) et.part[d] = have + __popcll(peers);
                if (on && have + before < et.lim[d]) {
                    const int64_t at = et.base[d] + have + before;
                    et.keys[at] = rows[c];
                    VT v;
                    set_pair(v, vals[2 * c], vals[2 * c + 1]);
                    et.vals[at] = v;
                }
            }
        }
        et.cursor += K * n_on;
        return;
    }
    if (!COUNT && on && et.cursor + K * n_on <= et.limit) {
        const int off = et.cursor + __popcll(mask & lanemask_lt());
#pragma unroll
        for (int c = 0; c < K; ++c) {
            et.keys[off + c * n_on] = rows[c];
            VT v;
            set_pair(v, vals[2 * c], vals[2 * c + 1]);
            et.vals[off + c * n_on] = v;
        }
    }
    et.cursor += K * n_on;
}

// The 8 corner rows of `cell` at level `lp` (corner c = cell + bits of c).
__device__ __forceinline__ void corner_rows(const LevelParams& lp, const uint32_t cell[3],
                                            uint32_t rows[8]) {
    cell_rows(lp, cell, rows);
}

// Corner i (0..3) of the face of a cell with bit `ax` fixed to `b`.
__device__ __forceinline__ int face_corner(int i, int ax, int b) {
    const int low = i & ((1 << ax) - 1);
    return low | (b << ax) | ((i >> ax) << (ax + 1));
}


// Position of cluster evaluation q (0..n-1) for this lane.
template <int CL>
__device__ __forceinline__ void cluster_point(const GeoArgs& a, const GeoSample& s, int q,
                                              float x[3]) {
    bool inb[3];
    if (CL == 0) {  // base, then FD 0..5
        if (q == 0) {
#pragma unroll
            for (int d = 0; d < 3; ++d) x[d] = s.x0[d];
        } else {
            offset_point(a, s.P, q - 1, x, inb);
        }
    } else {        // curvature 0..5
        offset_point(a, s.ps, q, x, inb);
    }
}

// Centre of a cluster: the base point (FD cluster) or the tangent-shifted
// point (curvature cluster; not itself evaluated), clamped and contracted as
// its offsets are.
template <int CL>
__device__ __forceinline__ void cluster_center(const GeoArgs& a, const GeoSample& s, float x[3]) {
#pragma unroll
    for (int d = 0; d < 3; ++d) x[d] = CL == 0 ? s.x0[d] : contract1(clampr(s.ps[d], a.r), a);
}

// Lane p + D (p - 1) of the lane's 16-lane row (a DPP row shift: no LDS
// traffic); lanes whose source lies outside the row get 0 and must not use it.
template <int D>
__device__ __forceinline__ uint32_t row_down(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x100 + D, 0xf, 0xf, false);
}
__device__ __forceinline__ uint32_t row_up1(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);
}

// One step of the suffix segmented sum over a 16-lane row: a lane that is
// not the end of its segment adds the accumulators of lane p + D (and takes
// over its end flag) when that lane is in the row.
template <int D>
__device__ __forceinline__ void merge_step(int p, float acc[16], bool& end) {
    float o[16];
#pragma unroll
    for (int c = 0; c < 16; ++c)
        o[c] = __int_as_float(
            __builtin_amdgcn_update_dpp(0, __float_as_int(acc[c]), 0x100 + D, 0xf, 0xf, false));
    const bool oend = row_down<D>((uint32_t)end) != 0u;
    if (!end && p + D < 16) {
#pragma unroll
        for (int c = 0; c < 16; ++c) acc[c] += o[c];
        end = oend;
    }
}

// Reduces the stored encode gradients of one cluster of evaluations into
// (row, value) pairs with a star-stencil merge.  Per level, B = the cell of
// the cluster centre; an evaluation whose cell C
//   * is B adds its 8 corner values into B's accumulators,
//   * is B +- e_axis (a face neighbour, the usual case of an offset of about
//     one cell) adds its 4 corners on the shared face into B's accumulators
//     and emits its 4 far corners,
//   * is anything else emits its 8 corners.
// B's accumulators are then merged with equal-B lanes of the same lane group
// (neighbouring samples of a ray) by a segmented sum in a fixed shuffle
// tree, and the segment head emits them.  Every sum has a fixed order.
// genc_base: LDS [eval][LJ][2][64] of this wave; the cluster's q-th
// evaluation is genc slot slot0 + q.  COUNT: positions only (no gradients);
// the cursor ends at the tile's pair count.
template <int LJ, int CL, bool COUNT, class VT>
__device__ __forceinline__ void flush_cluster(const GeoArgs& a, const LevelParams* lv,
                                              const GeoSample& s, bool valid, const float* genc_base,
                                              int n_q, int slot0, EmitTile<VT>& et) {
    const int l = lane_id(), p = l & 15, k = l >> 4;
#pragma unroll 1
    for (int j = 0; j < LJ; ++j) {
        const int level = k + 4 * j;
        const bool on = valid && level < a.active;
        LevelParams lp{};
        if (level < a.g.L) lp = lv[level];
        uint32_t B[3];
        {
            float xc[3];
            cluster_center<CL>(a, s, xc);
            const LevelPos pb = level_pos(lp, xc);
#pragma unroll
            for (int d = 0; d < 3; ++d) B[d] = pb.c[d];
        }
        bool touched = false;
        float acc[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) acc[q] = 0.0f;
        // encode gradients one evaluation ahead: each load pair is issued
        // before the previous evaluation's pair stores, so its latency
        // overlaps that evaluation's work instead of stalling every one
        // (carrying it across level slots too measured no better: more
        // registers)
        float g0n = 0.0f, g1n = 0.0f;
        if constexpr (!COUNT) {
            g0n = genc_base[(slot0 * LJ + j) * 128 + l];
            g1n = genc_base[(slot0 * LJ + j) * 128 + 64 + l];
        }
#pragma unroll 1
        for (int q = 0; q < n_q; ++q) {
            const float g0 = g0n, g1 = g1n;
            if constexpr (!COUNT) {
                if (q + 1 < n_q) {
                    g0n = genc_base[((slot0 + q + 1) * LJ + j) * 128 + l];
                    g1n = genc_base[((slot0 + q + 1) * LJ + j) * 128 + 64 + l];
                }
            }
            float x[3];
            cluster_point<CL>(a, s, q, x);
            const LevelPos pp = level_pos(lp, x);
            // classify C - B
            int ax = -1, sgn = 0, nz = 0;
            bool unit = true;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const int dd = (int)(pp.c[d] - B[d]);
                if (dd != 0) {
                    ++nz;
                    ax = d;
                    sgn = dd;
                    unit = unit && (dd == 1 || dd == -1);
                }
            }
            const bool same = nz == 0;
            const bool adj = nz == 1 && unit;
            const int axm = ax < 0 ? 0 : ax;
            const int far_bit = sgn > 0 ? 1 : 0;  // C's corners off B's face
            touched = touched || same || adj;
            // all 8 corners of C on their own
            {
                const bool em = on && !same && !adj;
                uint32_t rows[8];
                float v[16];
                if ((!COUNT || et.part) && em) corner_rows(lp, pp.c, rows);
                if (!COUNT && em) {
#pragma unroll
                    for (int c = 0; c < 8; ++c) {
                        const float w = corner_weight(pp, c);
                        v[2 * c] = w * g0;
                        v[2 * c + 1] = w * g1;
                    }
                }
                emit_pairs<8, COUNT, VT>(et, em, rows, v);
            }
            // the 4 far corners of a face neighbour
            {
                const bool em = on && adj;
                uint32_t rows[4];
                float v[8];
                if ((!COUNT || et.part) && em) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int c = face_corner(i, axm, far_bit);
                        rows[i] = lp.offset + grid_row(lp, pp.c[0] + (c & 1),
                                                       pp.c[1] + ((c >> 1) & 1),
                                                       pp.c[2] + ((c >> 2) & 1));
                    }
                }
                if (!COUNT && em) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int c = face_corner(i, axm, far_bit);
                        const float w = corner_weight(pp, c);
                        v[2 * i] = w * g0;
                        v[2 * i + 1] = w * g1;
                    }
                }
                emit_pairs<4, COUNT, VT>(et, em, rows, v);
            }
            // into B: corner t of B is corner t of C (same) or corner
            // t ^ (1 << ax) of C when t lies on the shared face (adj)
            if constexpr (!COUNT) {
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    const bool take = same || (adj && ((t >> axm) & 1) == far_bit);
                    const int c = same ? t : (t ^ (1 << axm));
                    const float w = corner_weight(pp, c);
                    if (take) {
                        acc[2 * t] = __fmaf_rn(w, g0, acc[2 * t]);
                        acc[2 * t + 1] = __fmaf_rn(w, g1, acc[2 * t + 1]);
                    }
                }
                // analytic mode, base evaluation (always in B): the gradient
                // term u * scale * grad(w_c) . g_x of each corner
                if (CL == 0 && a.analytic && q == 0) {
                    float ex[3][2];
#pragma unroll
                    for (int d = 0; d < 3; ++d)
#pragma unroll
                        for (int f = 0; f < 2; ++f)
                            ex[d][f] = genc_base[((a.n_evals + d) * LJ + j) * 128 + f * 64 + l];
#pragma unroll
                    for (int t = 0; t < 8; ++t) {
#pragma unroll
                        for (int f = 0; f < 2; ++f) {
                            const float dv = (corner_dweight(pp, t, 0) * ex[0][f] +
                                              corner_dweight(pp, t, 1) * ex[1][f]) +
                                             corner_dweight(pp, t, 2) * ex[2][f];
                            acc[2 * t + f] += dv;
                        }
                    }
                }
            }
        }
        // B: merge equal cells of neighbouring lanes (same lane group),
        // suffix segmented sum in a fixed tree, segment head emits.
        const uint32_t INV = 0xFFFFFFFFu;
        const bool live = on && touched;
        const uint32_t key0 = live ? B[0] : INV, key1 = live ? B[1] : INV,
                       key2 = live ? B[2] : INV;
        const uint32_t n0 = row_down<1>(key0), n1 = row_down<1>(key1), n2 = row_down<1>(key2);
        bool end = (p == 15) || !(n0 == key0 && n1 == key1 && n2 == key2) || key0 == INV;
        if constexpr (!COUNT) {
            merge_step<1>(p, acc, end);
            merge_step<2>(p, acc, end);
            merge_step<4>(p, acc, end);
            merge_step<8>(p, acc, end);
        }
        const uint32_t u0 = row_up1(key0), u1 = row_up1(key1), u2 = row_up1(key2);
        const bool head = live && (p == 0 || !(u0 == key0 && u1 == key1 && u2 == key2));
        uint32_t rows[8];
        if ((!COUNT || et.part) && head) corner_rows(lp, B, rows);
        emit_pairs<8, COUNT, VT>(et, head, rows, acc);
    }
}

// Rebuilds the per-sample state (grad, normal, tangent, curvature terms) of
// the lane's sample from the 12 stored values (6 FD values, or the analytic
// grad in the first 3, then the 6 curvature values).  n: the kernels' index
// (stored values), m: the caller's index (positions, directions).
__device__ __forceinline__ void geo_rebuild(const GeoArgs& a, const float* __restrict__ pts,
                                            const float* __restrict__ randn,
                                            const float* __restrict__ fdv, int64_t n, int64_t m,
                                            bool valid, GeoSample& s) {
#pragma unroll
    for (int d = 0; d < 3; ++d) s.P[d] = valid ? pts[3 * m + d] : 0.0f;
#pragma unroll
    for (int d = 0; d < 3; ++d) s.x0[d] = contract1(s.P[d], a);
    if (a.with_grad) {
        float g[3] = {0.0f, 0.0f, 0.0f};
        if (a.analytic) {
#pragma unroll
            for (int d = 0; d < 3; ++d) g[d] = valid ? fdv[n * 12 + d] : 0.0f;
        } else {
#pragma unroll
            for (int j = 0; j < 6; ++j) diff_pair(g, j, valid ? fdv[n * 12 + j] : 0.0f);
            finish_grad(a, g);
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) s.grad[d] = g[d];
        normalize3(s.grad, s.n);
    }
    if (a.with_laplace) {
        geo_tangent(a, randn, m, valid, s);
        float gs[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int j = 0; j < 6; ++j) diff_pair(gs, j, valid ? fdv[n * 12 + 6 + j] : 0.0f);
        finish_grad(a, gs);
#pragma unroll
        for (int d = 0; d < 3; ++d) s.grad_s[d] = gs[d];
        normalize3(s.grad_s, s.ns);
        s.dot = (s.n[0] * s.ns[0] + s.n[1] * s.ns[1]) + s.n[2] * s.ns[2];
    }
}

// Encode-gradient slots of a tile in HBM (written by the backward, read by
// the pair pass): slot 0 base, 1-6 FD, 7-12 curvature; a slot is [LJ][2][64]
// floats (level slot, feature, lane).
__host__ __device__ inline int64_t geo_genc_floats(int64_t tiles, int n_evals, int LJ) {
    return tiles * n_evals * LJ * 128;
}

// Pair pass.  COUNT: the number of (row, value) pairs each tile of 16
// samples emits, from positions alone.  Otherwise: the same decisions with
// the encode gradients, each tile's pairs written from offsets[tile] (never
// past counts[tile]; a different count sets status[1]).
// PART: the pairs are partitioned by their digit d = (row >> 9) & 255 as
// they are written: counts and offsets then hold kPartDigits entries per
// tile, digit-major ([d * n_tiles + tile]), and a tile's pairs of digit d go
// to [offsets[d n_tiles + tile], + counts[...]) in emission order.
template <int LJ, bool COUNT, class VT, bool PART>
__global__ void __launch_bounds__(256) k_geo_pairs(
    GeoArgs a, const float* __restrict__ pts, const float* __restrict__ randn, int64_t N,
    const float* __restrict__ fdv, const float* __restrict__ genc, uint32_t* __restrict__ keys,
    VT* __restrict__ vals, const int64_t* __restrict__ offsets, int64_t* __restrict__ counts,
    int64_t* __restrict__ status, int64_t capacity) {
    __shared__ LevelParams lv[AM_MAX_LEVELS];
    // per wave: running count, limit and first position of each digit
    __shared__ int32_t part_s[PART ? 4 : 1][PART ? kPartDigits : 1];
    __shared__ int32_t lim_s[PART && !COUNT ? 4 : 1][PART && !COUNT ? kPartDigits : 1];
    __shared__ int64_t base_s[PART && !COUNT ? 4 : 1][PART && !COUNT ? kPartDigits : 1];
#pragma unroll
    for (int q = 0; q < AM_MAX_LEVELS; ++q)
        if ((int)threadIdx.x == q && q < a.g.L) lv[q] = a.g.lv[q];
    __syncthreads();
    const int l = lane_id(), p = l & 15;
    const int nw = blockDim.x >> 6, wave = threadIdx.x >> 6;
    const int64_t n_tiles = (N + 15) / 16;
    for (int64_t tile = (int64_t)blockIdx.x * nw + wave; tile < n_tiles;
         tile += (int64_t)gridDim.x * nw) {
        const int64_t n = tile * 16 + p;
        const bool valid = n < N;
        GeoSample s;
        geo_rebuild(a, pts, randn, fdv, n, caller_row(a, n, valid), valid, s);
        EmitTile<VT> et{nullptr, nullptr, 0, 0, nullptr, nullptr, nullptr};
        const float* gt = nullptr;
        if constexpr (PART) {
            int32_t* part = part_s[wave];
            for (int d = l; d < kPartDigits; d += 64) {
                part[d] = 0;
                if constexpr (!COUNT) {
                    // a run past the arrays' capacity is cut there (and the
                    // consistency flag set): never a write out of range
                    const int64_t at = (int64_t)d * n_tiles + tile;
                    const int64_t b0 = offsets[at];
                    const int64_t room = capacity - b0;
                    const int64_t want = counts[at];
                    base_s[wave][d] = b0;
                    lim_s[wave][d] = (int32_t)(want <= room ? want : (room > 0 ? room : 0));
                    if (want > room) status[1] = 1;
                }
            }
            mlp::wave_sync();
            if constexpr (COUNT) {
                et.part = part;
            } else {
                et = EmitTile<VT>{keys, vals, 0, 0, part, lim_s[wave], base_s[wave]};
                gt = genc + tile * a.genc_slots * LJ * 128;
            }
        } else if constexpr (kAtomicPairs<VT>) {
            et = EmitTile<VT>{nullptr, vals, 0, 0, nullptr, nullptr, nullptr};
            gt = genc + tile * a.genc_slots * LJ * 128;
        } else if constexpr (!COUNT) {
            const int64_t off = offsets[tile];
            const int64_t room = capacity - off;
            int64_t lim = counts[tile];
            if (lim > room) {   // cut at the arrays' capacity, flagged
                lim = room > 0 ? room : 0;
                if (l == 0) status[1] = 1;
            }
            et = EmitTile<VT>{keys + off, vals + off, 0, (int)lim, nullptr, nullptr, nullptr};
            gt = genc + tile * a.genc_slots * LJ * 128;
        }
        if (a.with_laplace) flush_cluster<LJ, 1, COUNT>(a, lv, s, valid, gt, 6, a.curv0, et);
        flush_cluster<LJ, 0, COUNT>(a, lv, s, valid, gt, a.with_grad && !a.analytic ? 7 : 1, 0,
                                    et);
        if constexpr (PART) {
            mlp::wave_sync();
            bool bad = false;
            for (int d = l; d < kPartDigits; d += 64) {
                const int32_t got = et.part[d];
                if constexpr (COUNT) counts[(int64_t)d * n_tiles + tile] = got;
                else bad = bad || got != et.lim[d];
            }
            if constexpr (!COUNT)
                if (__ballot(bad) != 0 && l == 0) status[1] = 1;
            mlp::wave_sync();
        } else if constexpr (COUNT) {
            if (l == 0) counts[tile] = et.cursor;
        } else if constexpr (!kAtomicPairs<VT>) {
            if (l == 0 && et.cursor != et.limit) status[1] = 1;
        }
    }
}

template <int LJ, int NH, bool ST, int PR, bool AN>
__device__ __forceinline__ void geo_bwd_body(
    const GeoArgs& a, const float* __restrict__ pts,
    const float* __restrict__ randn, int64_t N, const void* __restrict__ table_v,
    const float* __restrict__ params, const float* __restrict__ fdv,
    const float* __restrict__ store,
    const float* __restrict__ g_sdf, const float* __restrict__ g_grad,
    const float* __restrict__ g_normal, const float* __restrict__ g_feature,
    const float* __restrict__ g_laplace, float* __restrict__ genc_g,
    float* __restrict__ partial) {
    extern __shared__ float sm[];
    __shared__ LevelParams lv[AM_MAX_LEVELS];
    geo_load_lds<LJ, NH>(a, params, sm, lv);
    __syncthreads();
    const TableT<PR>* table = reinterpret_cast<const TableT<PR>*>(table_v);
    const int l = lane_id(), p = l & 15, k = l >> 4;
    const int nw = blockDim.x >> 6, wave = threadIdx.x >> 6;
    float* scratch = sm + a.lds_floats + wave * (128 * mlp::kScratchLd);
    // accumulators of the feature rows of output tiles >= 1 (n_out > 16, two
    // hidden layers)
    const int ext_floats = NH == 2 ? (a.out_tiles - 1) * kOutExtFloats : 0;
    float* ext_acc = sm + geo_ext_base(a) + wave * ext_floats;
    for (int q = l; q < ext_floats; q += 64) ext_acc[q] = 0.0f;
    mlp::wave_sync();
    const int64_t n_tiles = (N + 15) / 16;
    const int fdim = a.n_out + 3;
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
        const int64_t m = caller_row(a, n, valid);
        // this tile's encode-gradient slots (0 base, then FD and curvature, or
        // curvature and the three analytic extra slots)
        float* genc = genc_g + tile * a.genc_slots * LJ * 128;
        GeoSample s;
        geo_rebuild(a, pts, randn, fdv, n, m, valid, s);
        // upstream gradients of this sample
        float gg[3] = {0, 0, 0}, gnrm[3] = {0, 0, 0};
        if (valid && g_grad)
#pragma unroll
            for (int d = 0; d < 3; ++d) gg[d] = g_grad[3 * m + d];
        if (valid && g_normal)
#pragma unroll
            for (int d = 0; d < 3; ++d) gnrm[d] = g_normal[3 * m + d];
        // output gradient of the base evaluation
        v4f go;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 4 * k + r;
            float v = 0.0f;
            if (valid && row < a.n_out) {
                if (g_feature) v = g_feature[m * fdim + row];
                if (row == 0 && g_sdf) v += g_sdf[m];
            }
            go[r] = v;
        }
        if constexpr (!AN) {   // base evaluation (genc slot 0)
            GeoEval<LJ, NH, false> e;
            geo_eval_b<LJ, NH, false, ST, mlp::kF32>(a, sm, lv, table, store, tile, n_tiles, 0, 0,
                                                     s.x0, e);
            const OutExt ext{g_feature, m, fdim, valid, ext_acc};
            geo_eval_bwd<LJ, NH, false, mlp::kF32>(a, sm, e, go, scratch, G, genc, nullptr,
                                                   a.out_tiles > 1 ? &ext : nullptr);
        }
        if (a.with_laplace) {
            const float gl = (valid && g_laplace) ? g_laplace[m] : 0.0f;
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
#pragma unroll 1
            for (int j = 0; j < 6; ++j) {
                const float ggs = pick3(g_gs, j >> 1);
                const float gf = (j & 1 ? -0.5f : 0.5f) * a.inv_eps * ggs;
                float x[3];
                bool inb[3];
                offset_point(a, s.ps, j, x, inb);
                GeoEval<LJ, NH, true> e;
                geo_eval_b<LJ, NH, true, ST, mlp::kF32>(a, sm, lv, table, store, tile, n_tiles,
                                                        a.curv0 + j, a.dx_curv0 + j, x, e);
                float gpos[3];
                geo_eval_bwd<LJ, NH, true, mlp::kF32>(a, sm, e, go_row0(gf), scratch, G,
                                                      genc + (a.curv0 + j) * LJ * 128, gpos);
#pragma unroll
                for (int d = 0; d < 3; ++d)
                    if (inb[d]) g_ps[d] += gpos[d] * a.inv2r;
            }
            float g_t[3], gcr[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) g_t[d] = g_ps[d] * a.eps;
            cross3(s.rn, g_t, gcr);
#pragma unroll
            for (int d = 0; d < 3; ++d) gnrm[d] += gcr[d];
        }
        if constexpr (AN) {
            // analytic mode: the base evaluation with the second-order terms
            // of the gradient output (with_grad holds in this mode)
            float gn2g[3];
            normalize3_bwd(s.grad, gnrm, gn2g);
            float gx[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) gx[d] = (gg[d] + gn2g[d]) * a.inv2r;
            GeoEval<LJ, NH, true> e;
            geo_eval_b<LJ, NH, true, ST, mlp::kF32>(a, sm, lv, table, store, tile, n_tiles, 0, 0,
                                                    s.x0, e);
            const OutExt ext{g_feature, m, fdim, valid, ext_acc};
            geo_eval_bwd_an<LJ, NH>(a, sm, lv, e, go, gx, scratch, G, genc,
                                    genc + a.n_evals * LJ * 128, a.out_tiles > 1 ? &ext : nullptr);
        } else if (a.with_grad) {
            float gn2g[3];
            normalize3_bwd(s.grad, gnrm, gn2g);
#pragma unroll
            for (int d = 0; d < 3; ++d) gg[d] += gn2g[d];
#pragma unroll 1
            for (int j = 0; j < 6; ++j) {
                const float gf = (j & 1 ? -0.5f : 0.5f) * a.inv_eps * pick3(gg, j >> 1);
                float x[3];
                bool inb[3];
                offset_point(a, s.P, j, x, inb);
                GeoEval<LJ, NH, false> e;
                geo_eval_b<LJ, NH, false, ST, mlp::kF32>(a, sm, lv, table, store, tile, n_tiles,
                                                         1 + j, 0, x, e);
                geo_eval_bwd<LJ, NH, false, mlp::kF32>(a, sm, e, go_row0(gf), scratch, G,
                                                       genc + (1 + j) * LJ * 128, nullptr);
            }
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
    // feature rows of output tiles >= 1
    mlp::wave_sync();
#pragma unroll 1
    for (int t = 1; t < (NH == 2 ? a.out_tiles : 1); ++t) {
        const float* acc = ext_acc + (t - 1) * kOutExtFloats;
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            const v4f dw = *reinterpret_cast<const v4f*>(acc + (it * 64 + l) * 4);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int o = 16 * t + 4 * k + r;
                if (o < a.n_out) out[a.pwo + o * 64 + 16 * it + i] = dw[r];
            }
        }
        if (i == 0)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int o = 16 * t + 4 * k + r;
                if (o < a.n_out) out[a.pbo + o] = acc[1024 + 4 * k + r];
            }
    }
}

#define AM_GEO_BWD_PARAMS                                                                  \
    GeoArgs a, const float* __restrict__ pts, const float* __restrict__ randn, int64_t N,      \
        const void* __restrict__ table_v, const float* __restrict__ params,                    \
        const float* __restrict__ fdv, const float* __restrict__ store,                        \
        const float* __restrict__ g_sdf, const float* __restrict__ g_grad,                     \
        const float* __restrict__ g_normal, const float* __restrict__ g_feature,               \
        const float* __restrict__ g_laplace, float* __restrict__ genc_g,                       \
        float* __restrict__ partial
#define AM_GEO_BWD_ARGS                                                                     \
    a, pts, randn, N, table_v, params, fdv, store, g_sdf, g_grad, g_normal, g_feature,          \
        g_laplace, genc_g, partial

template <int LJ, int NH, bool ST, int PR, bool AN>
__global__ void __launch_bounds__(256) k_geo_bwd(AM_GEO_BWD_PARAMS) {
    geo_bwd_body<LJ, NH, ST, PR, AN>(AM_GEO_BWD_ARGS);
}

// The 1-hidden-layer form at two waves per SIMD: its state fits the 256
// registers a lane has at that occupancy with a small spill, and the second
// wave hides the LDS and memory latency of the first.
template <int LJ, bool ST, int PR, bool AN>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_geo_bwd_o2(AM_GEO_BWD_PARAMS) {
    geo_bwd_body<LJ, 1, ST, PR, AN>(AM_GEO_BWD_ARGS);
}

// Sentinel keys and zero values in [status[0], bound): the unused tail of
// pair arrays carved for the worst-case bound.
template <class VT>
__global__ void k_pair_tail(uint32_t* __restrict__ keys, VT* __restrict__ vals,
                            const int64_t* __restrict__ status, int64_t bound, uint32_t sentinel) {
    const int64_t first = status[0];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = first + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < bound;
         i += stride) {
        keys[i] = sentinel;
        VT v;
        set_pair(v, 0.0f, 0.0f);
        vals[i] = v;
    }
}

// status[0] = number of pairs (offset + count of the last entry: the last
// tile, or the last digit of the last tile); status[1] = 0, set by the
// backward on a count mismatch.
__global__ void k_pair_total(const int64_t* __restrict__ offsets, const int64_t* __restrict__ counts,
                             int64_t tiles, int64_t* __restrict__ status) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        status[0] = offsets[tiles - 1] + counts[tiles - 1];
        status[1] = 0;
    }
}

// ------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------
static int make_geo_args(const am_grid_desc* g, const am_mlp_desc* m, const am_sdf_desc* s,
                         GeoArgs* out, int* LJ) {
    AM_CHECK_ARG(g && m && s, "NULL descriptor");
    if (int rc = check_mlp_common(m, kGeoMaxOut)) return rc;
    if (m->n_out > 16 && m->n_hidden_layers != 2) {
        set_error("an SDF MLP with more than 16 outputs (got %d) takes two hidden layers",
                  m->n_out);
        return AM_ERR_UNSUPPORTED;
    }
    AM_CHECK_ARG(g->n_features == 2, "fused geometry needs n_features = 2 (got %d)", g->n_features);
    AM_CHECK_ARG(g->n_levels >= 1 && g->n_levels <= 16, "fused geometry needs <= 16 levels");
    AM_CHECK_ARG(m->n_in == 3 + 2 * g->n_levels, "SDF MLP n_in %d != 3 + L*F", m->n_in);
    AM_CHECK_ARG(s->active_levels >= 0 && s->active_levels <= g->n_levels, "active_levels %d",
                 s->active_levels);
    AM_CHECK_ARG(s->radius > 0.0f, "radius <= 0");
    AM_CHECK_ARG(s->grad_type == AM_GRAD_FD || s->grad_type == AM_GRAD_ANALYTIC, "grad_type %d",
                 s->grad_type);
    const bool analytic = s->with_grad && s->grad_type == AM_GRAD_ANALYTIC;
    AM_CHECK_ARG(!(s->with_laplace || (s->with_grad && !analytic)) || s->fd_eps > 0.0f,
                 "fd_eps <= 0");
    AM_CHECK_ARG(!s->with_laplace || s->with_grad, "laplace needs with_grad");
    AM_CHECK_ARG(s->precision == AM_PREC_FP32 || s->precision == AM_PREC_MIXED,
                 "precision %d", s->precision);
    AM_CHECK_ARG(m->precision == AM_PREC_FP32,
                 "the SDF MLP runs in fp32 in every form (got precision %d): central "
                 "differences amplify per-evaluation rounding by 1/eps", m->precision);
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
    if (m->n_hidden_layers == 2) { a.w2 = o; o += 64 * kWLd; a.b2 = o; o += 64; }
    a.out_tiles = (m->n_out + 15) / 16;
    a.wo = o; o += a.out_tiles * 16 * kWLd;
    a.bo = o; o += a.out_tiles * 16;
    a.lds_floats = o;
    layout_params(m->n_in, m->n_out, m->n_hidden_layers, &a.pw1, &a.pb1, &a.pw2, &a.pb2, &a.pwo,
                  &a.pbo, &a.n_params);
    a.analytic = analytic ? 1 : 0;
    a.n_evals = 1 + (s->with_grad && !analytic ? 6 : 0) + (s->with_laplace ? 6 : 0);
    a.curv0 = analytic ? 1 : 7;
    a.dx_curv0 = analytic ? 1 : 0;
    a.n_dx = (analytic ? 1 : 0) + (s->with_laplace ? 6 : 0);
    a.genc_slots = a.n_evals + (analytic ? 3 : 0);
    a.prec = s->precision;
    *out = a;
    *LJ = lj;
    return AM_OK;
}

template <int LJ, int NH, bool ST, int PR>
static int geo_fwd_launch_t(const GeoArgs& a, const float* pts, const float* randn, int64_t n,
                            const void* table, const float* params, float* sdf, float* grad,
                            float* normal, float* feature, float* laplace, float* fdv,
                            float* store, hipStream_t s) {
    const size_t lds = (size_t)a.lds_floats * sizeof(float);
    auto kern = a.analytic ? k_geo_fwd<LJ, NH, ST, PR, true> : k_geo_fwd<LJ, NH, ST, PR, false>;
    AM_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
    {
        KernelTimer kt_("geo_fwd", s);
        hipLaunchKernelGGL(kern, dim3(fwd_blocks(n)), dim3(256), lds, s, a, pts, randn, n, table,
                           params, sdf, grad, normal, feature, laplace, fdv, store);
    }
    AM_LAUNCH_CHECK();
    return AM_OK;
}

template <int LJ, int NH>
static int geo_fwd_launch(const GeoArgs& a, const float* pts, const float* randn, int64_t n,
                          const void* table, const float* params, float* sdf, float* grad,
                          float* normal, float* feature, float* laplace, float* fdv, float* store,
                          hipStream_t s) {
    if (a.prec == AM_PREC_MIXED) {
        if (store)
            return geo_fwd_launch_t<LJ, NH, true, mlp::kBF16>(a, pts, randn, n, table, params, sdf,
                                                              grad, normal, feature, laplace, fdv,
                                                              store, s);
        return geo_fwd_launch_t<LJ, NH, false, mlp::kBF16>(a, pts, randn, n, table, params, sdf,
                                                           grad, normal, feature, laplace, fdv,
                                                           nullptr, s);
    }
    if (store)
        return geo_fwd_launch_t<LJ, NH, true, mlp::kF32>(a, pts, randn, n, table, params, sdf,
                                                         grad, normal, feature, laplace, fdv,
                                                         store, s);
    return geo_fwd_launch_t<LJ, NH, false, mlp::kF32>(a, pts, randn, n, table, params, sdf, grad,
                                                      normal, feature, laplace, fdv, nullptr, s);
}

template <int LJ>
static int geo_count_launch(const GeoArgs& a, const float* pts, const float* randn, int64_t n,
                            const float* fdv, int64_t* counts, bool part, hipStream_t s) {
    KernelTimer kt_("pair_count", s);
    if (part)
        hipLaunchKernelGGL((k_geo_pairs<LJ, true, Vals<2>, true>), dim3(count_blocks(n)), dim3(256),
                           0, s, a, pts, randn, n, fdv, nullptr, nullptr, nullptr, nullptr, counts,
                           nullptr, (int64_t)0);
    else
        hipLaunchKernelGGL((k_geo_pairs<LJ, true, Vals<2>, false>), dim3(count_blocks(n)),
                           dim3(256), 0, s, a, pts, randn, n, fdv, nullptr, nullptr, nullptr,
                           nullptr, counts, nullptr, (int64_t)0);
    AM_LAUNCH_CHECK();
    return AM_OK;
}

// Partitioned pair emission applies when the table reduce sorts key bits
// [9, end) (9 < end <= 25) and the "table_prepartition" flag is on (off by
// default: it measured slower, see kPartDigits).
static bool geo_prepartition(const am_grid_desc* g) {
    return table_bucket_bits(table_key_bits(g->n_entries)) == kPartShift &&
           flag_value("table_prepartition", 0) != 0;
}

// The backward's workspace before the pair arrays: the pair count is only
// known after the count pass, so the sort buffers are carved afterwards
// from the rest (sized for the worst case by the size query).
struct GeoBwdWs {
    float* fdv;
    float* sdf_tmp;
    float* genc;
    float* partial;
    float* ctmp;
    int64_t* counts;
    int64_t* offsets;
    int64_t entries;   // count / offset entries (tiles, or tiles x digits)
    int64_t* status;
    void* scan_tmp;
    size_t scan_bytes;
};

static void carve_geo_bwd_head(WsCarver& c, const GeoArgs& a, int LJ, int64_t n, bool need_fdv,
                               bool part, GeoBwdWs* w) {
    const int waves = bwd_blocks(n) * 4;
    const int64_t tiles = (n + 15) / 16;
    const int64_t entries = tiles * (part ? kPartDigits : 1);   // count / offset entries
    w->fdv = need_fdv ? c.take<float>((size_t)n * 12 + 1) : nullptr;
    w->sdf_tmp = need_fdv ? c.take<float>((size_t)n + 1) : nullptr;
    w->genc = c.take<float>((size_t)geo_genc_floats(tiles, a.genc_slots, LJ) + 1);
    w->partial = c.take<float>((size_t)waves * a.n_params);
    w->ctmp = c.take<float>((size_t)colsum_tmp_floats(waves, a.n_params) + 1);
    w->counts = c.take<int64_t>((size_t)entries + 1);
    w->offsets = c.take<int64_t>((size_t)entries + 1);
    w->entries = entries;
    w->status = c.take<int64_t>(2);
    size_t sb = 0;
    rocprim::exclusive_scan(nullptr, sb, (const int64_t*)nullptr, (int64_t*)nullptr, (int64_t)0,
                            (size_t)(entries > 0 ? entries : 1), rocprim::plus<int64_t>(), 0,
                            false);
    w->scan_bytes = sb;
    w->scan_tmp = c.take<char>(sb + 256);
}

// Most pairs a tile can emit, per lane and level slot.  FD cluster: the
// base point is always in B, so 6 x 8 + 8 for B = 56; analytic or without
// gradient: the base alone, 8.  Curvature cluster (its centre is not
// evaluated): 6 x 8 = 48 with B untouched; with one point in a face
// neighbour of B, 5 x 8 + 4 far corners + 8 for B = 52 (the most); more
// points in or beside B emit less.
static int64_t geo_max_pairs(const GeoArgs& a, int LJ, int64_t n) {
    const int64_t per_slot = (a.with_grad && !a.analytic ? 56 : 8) + (a.with_laplace ? 52 : 0);
    return (n + 15) / 16 * (per_slot * LJ * 64);
}

// Consistency flags of recent backward calls.  The emit pass flags a tile
// whose pairs differ from the count pass's; instead of a stream sync per
// call, each call copies its flag into a pinned host slot behind an event,
// and later calls (after their own pair-count sync, cheaply) and
// am_sdf_bwd_check read the slots whose copies have landed.  Every emitted
// pair stays inside its tile's reserved range and the sort and reduce only
// index rows below n_entries and items below the count whatever the keys
// hold, so a mismatch gives a wrong dtable, never a stray access.
namespace {
struct FlagSlot {
    hipEvent_t ev = nullptr;
    int64_t* flag = nullptr;   // pinned host memory
    bool armed = false;
};
constexpr int kFlagSlots = 16;
std::mutex g_flag_mu;
FlagSlot g_flag_slots[kFlagSlots];
int g_flag_next = 0;
}  // namespace

static int flag_result(FlagSlot& f) {
    f.armed = false;
    if (*f.flag != 0) {
        set_error("table-gradient pair count mismatch between the count and emit passes of an "
                  "earlier am_sdf_bwd call");
        return AM_ERR_INTERNAL;
    }
    return AM_OK;
}

// Reads the armed slots: those whose copy has landed, or all (wait).
static int collect_flags(bool wait) {
    std::lock_guard<std::mutex> lk(g_flag_mu);
    int rc = AM_OK;
    for (FlagSlot& f : g_flag_slots) {
        if (!f.armed) continue;
        const hipError_t e = wait ? hipEventSynchronize(f.ev) : hipEventQuery(f.ev);
        if (e == hipErrorNotReady) continue;
        if (e != hipSuccess) {
            set_error("consistency flag event failed: %s", hipGetErrorString(e));
            return AM_ERR_HIP;
        }
        if (int r = flag_result(f)) rc = r;
    }
    return rc;
}

static int arm_flag(const int64_t* dev_flag, hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_flag_mu);
    FlagSlot& f = g_flag_slots[g_flag_next];
    g_flag_next = (g_flag_next + 1) % kFlagSlots;
    if (f.ev == nullptr) {
        AM_HIP(hipEventCreateWithFlags(&f.ev, hipEventDisableTiming));
        AM_HIP(hipHostMalloc((void**)&f.flag, sizeof(int64_t), hipHostMallocDefault));
    } else if (f.armed) {   // the ring wrapped onto a pending slot
        AM_HIP(hipEventSynchronize(f.ev));
        if (int rc = flag_result(f)) return rc;
    }
    AM_HIP(hipMemcpyAsync(f.flag, dev_flag, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    AM_HIP(hipEventRecord(f.ev, s));
    f.armed = true;
    return AM_OK;
}

template <int LJ, int NH, int PR>
static int geo_bwd_launch_t(const GeoArgs& a, const am_grid_desc* g, const float* pts,
                            const float* randn, int64_t n, const void* table, const float* params,
                            const float* fdv_in,
                          const float* store, const float* g_sdf, const float* g_grad,
                          const float* g_normal, const float* g_feature, const float* g_laplace,
                          float* dtable, float* dparams, void* ws, size_t ws_bytes, hipStream_t s) {
    const int blocks = bwd_blocks(n);
    const int waves = blocks * 4;
    const int64_t tiles = (n + 15) / 16;
    const bool need_fdv = a.with_grad && fdv_in == nullptr;
    AM_CHECK_ARG(!store || !need_fdv, "a stored encoding needs the stored FD values too");
    WsCarver c(ws, ws_bytes);
    GeoBwdWs w;
    const bool part = geo_prepartition(g);
    carve_geo_bwd_head(c, a, LJ, n, need_fdv, part, &w);
    if (!c.ok()) return workspace_short(c.used, ws_bytes);
    const float* fdv = fdv_in;
    if (need_fdv) {  // recompute the stored SDF values with a forward pass
        if (int rc = geo_fwd_launch<LJ, NH>(a, pts, randn, n, table, params, w.sdf_tmp, nullptr,
                                            nullptr, nullptr, nullptr, w.fdv, nullptr, s))
            return rc;
        fdv = w.fdv;
    }
    // the backward proper: MLP parameter partials and the encode gradients
    auto launch_bwd = [&]() -> int {
        const size_t lds =
            (size_t)(geo_ext_base(a) + 4 * (a.out_tiles - 1) * kOutExtFloats) * sizeof(float);
        AM_CHECK_ARG(lds <= 160 * 1024, "geometry backward needs %zu bytes of LDS", lds);
        auto kern = k_geo_bwd_o2<LJ, true, PR, false>;   // the type of every form's kernel
        if constexpr (NH == 1) {
            kern = store ? k_geo_bwd_o2<LJ, true, PR, false> : k_geo_bwd_o2<LJ, false, PR, false>;
            if (a.analytic)
                kern = store ? k_geo_bwd_o2<LJ, true, PR, true> : k_geo_bwd_o2<LJ, false, PR, true>;
        } else {
            kern = store ? k_geo_bwd<LJ, NH, true, PR, false> : k_geo_bwd<LJ, NH, false, PR, false>;
            if (a.analytic)
                kern = store ? k_geo_bwd<LJ, NH, true, PR, true> : k_geo_bwd<LJ, NH, false, PR, true>;
        }
        AM_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
        {
            KernelTimer kt_("geo_bwd", s);
            hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), lds, s, a, pts, randn, n, table,
                               params, fdv, store, g_sdf, g_grad, g_normal, g_feature, g_laplace,
                               w.genc, w.partial);
        }
        AM_LAUNCH_CHECK();
        return AM_OK;
    };
    if (flag_value("table_atomic", 0) != 0) {
        // atomic form: the pairs are added into the zeroed table gradient
        if (int rc = fill_f32(dtable, (int64_t)g->n_entries * 2, 0.0f, s)) return rc;
        if (int rc = launch_bwd()) return rc;
        {
            KernelTimer kt_("pair_atomic", s);
            hipLaunchKernelGGL((k_geo_pairs<LJ, false, AtomicRow, false>), dim3(pair_blocks(n)),
                               dim3(256), 0, s, a, pts, randn, n, fdv, w.genc, nullptr,
                               reinterpret_cast<AtomicRow*>(dtable), nullptr, nullptr, nullptr,
                               (int64_t)0);
        }
        AM_LAUNCH_CHECK();
        return colsum_f32(w.partial, waves, a.n_params, dparams, w.ctmp, s);
    }
    // 1. pairs per tile, 2. exclusive prefix sum -> each tile's first pair
    if (int rc = geo_count_launch<LJ>(a, pts, randn, n, fdv, w.counts, part, s)) return rc;
    size_t sb = w.scan_bytes;
    hipError_t e = rocprim::exclusive_scan(w.scan_tmp, sb, w.counts, w.offsets, (int64_t)0,
                                           (size_t)w.entries, rocprim::plus<int64_t>(), s, false);
    if (e != hipSuccess) {
        set_error("tile-count scan failed: %s", hipGetErrorString(e));
        return AM_ERR_HIP;
    }
    hipLaunchKernelGGL(k_pair_total, dim3(1), dim3(64), 0, s, w.offsets, w.counts, w.entries,
                       w.status);
    AM_LAUNCH_CHECK();
    // 3. the pair arrays.  Sync-free form (flag "sdf_bwd_sync_free"): arrays
    // of the worst-case pair bound, no host read; the tail past the
    // device-side count is filled with sentinel keys (they sort last and are
    // skipped) and the sort covers the bound.  Otherwise one host read of the
    // pair count carves the arrays for exactly that many pairs.  A workspace
    // too small is rejected here, before any output is written, with
    // am_last_workspace_need() = what this call needs.
    const int64_t max_pairs = geo_max_pairs(a, LJ, n);
    const bool bounded = flag_value("sdf_bwd_sync_free", 0) != 0;
    int64_t st[2] = {0, 0};
    if (!bounded) {
        AM_HIP(hipMemcpyAsync(st, w.status, sizeof(st), hipMemcpyDeviceToHost, s));
        AM_HIP(hipStreamSynchronize(s));
        timer_units("host_syncs", 1);
    }
    if (int rc = collect_flags(false)) return rc;
    TableReduceWs tw;
    if (int rc = table_reduce_carve(c, 2, bounded ? max_pairs : st[0], g->n_entries, &tw,
                                    PR == mlp::kBF16, part))
        return rc;
    if (!c.ok()) return workspace_short(c.used, ws_bytes);
    tw.m = bounded ? max_pairs : st[0];   // the arrays' capacity (and the sort's size)
    if (bounded && max_pairs > 0) {
        const int64_t want = ceil_div(max_pairs, 256);
        hipLaunchKernelGGL((k_pair_tail<PairV<PR>>), dim3((unsigned)(want < 16384 ? want : 16384)),
                           dim3(256), 0, s, tw.keys[0], reinterpret_cast<PairV<PR>*>(tw.vals[0]),
                           w.status, max_pairs, table_sentinel(tw.end_bit));
        AM_LAUNCH_CHECK();
    }
    if (int rc = launch_bwd()) return rc;
    {
        KernelTimer kt_("pair_emit", s);
        if (part)
            hipLaunchKernelGGL((k_geo_pairs<LJ, false, PairV<PR>, true>), dim3(pair_blocks(n)), dim3(256),
                               0, s, a, pts, randn, n, fdv, w.genc, tw.keys[0],
                               reinterpret_cast<PairV<PR>*>(tw.vals[0]), w.offsets, w.counts,
                               w.status, tw.m);
        else
            hipLaunchKernelGGL((k_geo_pairs<LJ, false, PairV<PR>, false>), dim3(pair_blocks(n)),
                               dim3(256), 0, s, a, pts, randn, n, fdv, w.genc, tw.keys[0],
                               reinterpret_cast<PairV<PR>*>(tw.vals[0]), w.offsets, w.counts,
                               w.status, tw.m);
    }
    AM_LAUNCH_CHECK();
    if (int rc = colsum_f32(w.partial, waves, a.n_params, dparams, w.ctmp, s)) return rc;
    if (!bounded) timer_units("table_pairs", tw.m);
    if (int rc = table_reduce_run(tw, 2, g->n_entries, dtable, s)) return rc;
    return arm_flag(w.status + 1, s);
}

template <int LJ, int NH>
static int geo_bwd_launch(const GeoArgs& a, const am_grid_desc* g, const am_mlp_desc* m,
                          const am_sdf_desc* sd, const float* pts, const float* randn, int64_t n,
                          const void* table, const float* params, const float* fdv_in,
                          const float* store, const float* g_sdf, const float* g_grad,
                          const float* g_normal, const float* g_feature, const float* g_laplace,
                          float* dtable, float* dparams, void* ws, size_t ws_bytes, hipStream_t s) {
    if (a.prec == AM_PREC_MIXED)
        return geo_bwd_launch_t<LJ, NH, mlp::kBF16>(a, g, pts, randn, n, table, params, fdv_in,
                                                    store, g_sdf, g_grad, g_normal, g_feature,
                                                    g_laplace, dtable, dparams, ws, ws_bytes, s);
    return geo_bwd_launch_t<LJ, NH, mlp::kF32>(a, g, pts, randn, n, table, params, fdv_in, store,
                                               g_sdf, g_grad, g_normal, g_feature, g_laplace,
                                               dtable, dparams, ws, ws_bytes, s);
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

extern "C" int64_t am_sdf_store_floats(const am_grid_desc* g, const am_mlp_desc* m,
                                       const am_sdf_desc* sd, int64_t n) {
    GeoArgs a;
    int LJ;
    if (make_geo_args(g, m, sd, &a, &LJ) || n < 0) return -1;
    const int64_t tiles = (n + 15) / 16;
    return geo_enc_floats(tiles, a.n_evals, LJ) + geo_dx_floats(tiles, a.n_dx, LJ);
}

static int sdf_fwd(const am_grid_desc* g, const am_mlp_desc* m, const am_sdf_desc* sd,
                   const int32_t* order, const float* pts, const float* randn, int64_t n,
                   const void* table, const float* params, float* sdf, float* grad,
                   float* normal, float* feature, float* laplace, float* fd_values,
                   float* enc_store, void* stream) {
    GeoArgs a;
    int LJ;
    if (int rc = make_geo_args(g, m, sd, &a, &LJ)) return rc;
    a.order = order;
    AM_CHECK_ARG(n >= 0, "n < 0");
    if (n == 0) return AM_OK;
    AM_CHECK_ARG(pts && table && params && sdf, "NULL buffer");
    AM_CHECK_ARG(!sd->with_laplace || randn, "laplace needs randn");
    hipStream_t s = to_stream(stream);
    AM_GEO_DISPATCH(geo_fwd_launch, a, pts, randn, n, table, params, sdf, grad, normal, feature,
                    laplace, fd_values, enc_store, s);
}

extern "C" int am_sdf_fwd(const am_grid_desc* g, const am_mlp_desc* m, const am_sdf_desc* sd,
                          const float* pts, const float* randn, int64_t n, const void* table,
                          const float* params, float* sdf, float* grad, float* normal,
                          float* feature, float* laplace, float* fd_values, float* enc_store,
                          void* stream) {
    return sdf_fwd(g, m, sd, nullptr, pts, randn, n, table, params, sdf, grad, normal, feature,
                   laplace, fd_values, enc_store, stream);
}

extern "C" int am_sdf_fwd_ordered(const am_grid_desc* g, const am_mlp_desc* m,
                                  const am_sdf_desc* sd, const int32_t* order, const float* pts,
                                  const float* randn, int64_t n, const void* table,
                                  const float* params, float* sdf, float* grad, float* normal,
                                  float* feature, float* laplace, float* fd_values,
                                  float* enc_store, void* stream) {
    AM_CHECK_ARG(order || n == 0, "NULL order");
    return sdf_fwd(g, m, sd, order, pts, randn, n, table, params, sdf, grad, normal, feature,
                   laplace, fd_values, enc_store, stream);
}

extern "C" size_t am_sdf_bwd_workspace_size(const am_grid_desc* g, const am_mlp_desc* m,
                                            const am_sdf_desc* sd, int64_t n) {
    GeoArgs a;
    int LJ;
    if (make_geo_args(g, m, sd, &a, &LJ) || n < 0) return 0;
    WsCarver c(nullptr, 0);
    GeoBwdWs w;
    const bool part = geo_prepartition(g);
    carve_geo_bwd_head(c, a, LJ, n, true, part, &w);
    TableReduceWs tw;
    if (table_reduce_carve(c, 2, geo_max_pairs(a, LJ, n), g->n_entries, &tw, false, part))
        return 0;
    return c.used + 256;
}

extern "C" int am_sdf_bwd_check(void) { return collect_flags(true); }

static int sdf_bwd(const am_grid_desc* g, const am_mlp_desc* m, const am_sdf_desc* sd,
                   const int32_t* order, const float* pts, const float* randn, int64_t n,
                   const void* table, const float* params, const float* fd_values,
                   const float* enc_store, const float* g_sdf, const float* g_grad,
                   const float* g_normal, const float* g_feature, const float* g_laplace,
                   float* dtable, float* dparams, void* workspace, size_t workspace_bytes,
                   void* stream) {
    GeoArgs a;
    int LJ;
    if (int rc = make_geo_args(g, m, sd, &a, &LJ)) return rc;
    a.order = order;
    AM_CHECK_ARG(n >= 0, "n < 0");
    AM_CHECK_ARG(table && params && dtable && dparams && workspace, "NULL buffer");
    hipStream_t s = to_stream(stream);
    if (n == 0) {
        if (int rc = fill_f32(dtable, (int64_t)g->n_entries * 2, 0.0f, s)) return rc;
        return fill_f32(dparams, a.n_params, 0.0f, s);
    }
    AM_CHECK_ARG(pts, "NULL pts");
    AM_CHECK_ARG(!sd->with_laplace || randn, "laplace needs randn");
    AM_GEO_DISPATCH(geo_bwd_launch, a, g, m, sd, pts, randn, n, table, params, fd_values,
                    enc_store, g_sdf, g_grad, g_normal, g_feature, g_laplace, dtable, dparams,
                    workspace, workspace_bytes, s);
}

extern "C" int am_sdf_bwd(const am_grid_desc* g, const am_mlp_desc* m, const am_sdf_desc* sd,
                          const float* pts, const float* randn, int64_t n, const void* table,
                          const float* params, const float* fd_values, const float* enc_store,
                          const float* g_sdf, const float* g_grad, const float* g_normal,
                          const float* g_feature, const float* g_laplace, float* dtable,
                          float* dparams, void* workspace, size_t workspace_bytes, void* stream) {
    return sdf_bwd(g, m, sd, nullptr, pts, randn, n, table, params, fd_values, enc_store, g_sdf,
                   g_grad, g_normal, g_feature, g_laplace, dtable, dparams, workspace,
                   workspace_bytes, stream);
}

extern "C" int am_sdf_bwd_ordered(const am_grid_desc* g, const am_mlp_desc* m,
                                  const am_sdf_desc* sd, const int32_t* order, const float* pts,
                                  const float* randn, int64_t n, const void* table,
                                  const float* params, const float* fd_values,
                                  const float* enc_store, const float* g_sdf, const float* g_grad,
                                  const float* g_normal, const float* g_feature,
                                  const float* g_laplace, float* dtable, float* dparams,
                                  void* workspace, size_t workspace_bytes, void* stream) {
    AM_CHECK_ARG(order || n == 0, "NULL order");
    return sdf_bwd(g, m, sd, order, pts, randn, n, table, params, fd_values, enc_store, g_sdf,
                   g_grad, g_normal, g_feature, g_laplace, dtable, dparams, workspace,
                   workspace_bytes, stream);
}
