// Wavefront-resident MLP engine on CDNA4 fp32 MFMA (v_mfma_f32_16x16x4f32).
//
// A wave evaluates a tile of 16 points.  Lane l serves point p = l % 16 and
// lane group k = l / 16.  For D = A(16x4) * B(4x16) the operand layouts are
// A[l%16][l/16], B[l/16][l%16], D[4*(l/16)+r][l%16] (r = 0..3); a layer
// computes D[out][p] = W[out][in] * X[in][p], so points are the N dimension.
//
// Row space: the B fragment of K-step s in lane group k holds input row
// krow(s,k) = 16*(s/4) + 4*k + s%4.  That is exactly the row a D tile t,
// register r puts in lane group k (t = s/4, r = s%4), so a layer's output
// registers are the next layer's B fragments as they stand, and an input
// gradient W^T*G comes out in the same lanes that hold the inputs.  Weight
// matrices sit in LDS as [out rows][ld] with columns in row space.
//
// Fragments are read through FragF (plain float arrays) or FragV (D tiles,
// fragment s = tile s/4, register s%4) so register arrays are never
// reinterpreted through pointers (which would push them to scratch).
//
// BF16 form (PR = kBF16): v_mfma_f32_16x16x16_bf16 takes in lane group k
// the 4 operand values k' = 4k..4k+3 of a 16-deep K step, and its D layout
// is the fp32 one.  Four consecutive fp32 K-steps 4u..4u+3 of a lane group
// therefore become one bf16 step with the values packed in step order: the
// same row space, weights and fragments, the products summed in fp32.
// Operands are rounded to bf16 (round to nearest even, v_cvt_pk_bf16_f32).
#pragma once
#include "am_common.h"

namespace am {
namespace mlp {

__device__ __forceinline__ int krow(int s, int k) { return ((s >> 2) << 4) + (k << 2) + (s & 3); }

constexpr int kF32 = 0;
constexpr int kBF16 = 1;

typedef short s4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf4v __attribute__((ext_vector_type(4)));

// One 16-deep bf16 step: a[q], b[q] are this lane's operands of fp32 K-step
// 4u + q.
__device__ __forceinline__ v4f mfma_bf16_step(v4f a, v4f b, v4f c) {
    const bf4v ab = __builtin_convertvector(a, bf4v);
    const bf4v bb = __builtin_convertvector(b, bf4v);
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(s4v, ab),
                                                     __builtin_bit_cast(s4v, bb), c, 0, 0, 0);
}

// Value as the bf16 MFMA sees it (round to nearest even).
__device__ __forceinline__ float round_bf16(float x) { return (float)(__bf16)x; }

struct FragF {
    const float* p;
    __device__ __forceinline__ float operator[](int s) const { return p[s]; }
};
struct FragV {
    const v4f* p;
    __device__ __forceinline__ float operator[](int s) const { return p[s >> 2][s & 3]; }
};

// Orders LDS traffic of one wave (lanes exchange data through LDS).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- activations (torch semantics) ----
// log1p(t) for t in [0, 1]: t * ln(u) / (u - 1) with u = fl(1 + t), which
// cancels the rounding of u (t itself when u == 1); hardware log2/rcp.  The
// method: D. Goldberg, "What every computer scientist should know about
// floating-point arithmetic", ACM Computing Surveys 23(1), 1991, Theorem 4.
__device__ __forceinline__ float log1p_unit(float t) {
    const float u = 1.0f + t;
    const float d = u - 1.0f;
    const float l = __builtin_amdgcn_logf(u) * 0.693147180559945309f;
    return d == 0.0f ? t : t * l * __builtin_amdgcn_rcpf(d);
}

// torch softplus(z, beta, threshold 20) = bz > 20 ? z : log1p(exp(bz)) / beta
// with bz = beta z, in the overflow-free form (max(bz, 0) + log1p(t)) / beta,
// t = exp(-|bz|) in (0, 1]; its derivative sigmoid(bz) is 1/(1+t) (bz >= 0)
// or t/(1+t) from the same t.
__device__ __forceinline__ float softplus_t(float bz) {
    return __builtin_amdgcn_exp2f(-fabsf(bz) * 1.44269504088896341f);
}

__device__ __forceinline__ float softplus_fwd(float z, float beta) {
    const float bz = z * beta;
    const float t = softplus_t(bz);
    return bz > 20.0f ? z : (fmaxf(bz, 0.0f) + log1p_unit(t)) * (1.0f / beta);
}

__device__ __forceinline__ void softplus_fb(float z, float beta, float& h, float& dh) {
    const float bz = z * beta;
    const float t = softplus_t(bz);
    const float r = __builtin_amdgcn_rcpf(1.0f + t);
    const bool big = bz > 20.0f;
    h = big ? z : (fmaxf(bz, 0.0f) + log1p_unit(t)) * (1.0f / beta);
    dh = big ? 1.0f : (bz >= 0.0f ? r : t * r);
}

__device__ __forceinline__ float act_fwd(int act, float z, float beta) {
    if (act == AM_ACT_RELU) return z > 0.0f ? z : 0.0f;
    if (act == AM_ACT_SOFTPLUS) return softplus_fwd(z, beta);
    if (act == AM_ACT_SIGMOID) return 1.0f / (1.0f + expf(-z));
    return z;
}

// d act / d z evaluated at pre-activation z.
__device__ __forceinline__ float act_bwd(int act, float z, float beta) {
    if (act == AM_ACT_RELU) return z > 0.0f ? 1.0f : 0.0f;
    if (act == AM_ACT_SOFTPLUS) {
        float h, dh;
        softplus_fb(z, beta, h, dh);
        return dh;
    }
    if (act == AM_ACT_SIGMOID) {
        const float y = 1.0f / (1.0f + expf(-z));
        return (1.0f - y) * y;
    }
    return 1.0f;
}

// Second derivative of the activation at z given its first derivative dh
// there (the analytic-gradient backward): softplus beta * s * (1 - s) with
// s = sigmoid(beta z) = dh (0 above torch's threshold, where dh = 1), 0 for
// relu and identity, sigmoid y (1 - y)(1 - 2y).
__device__ __forceinline__ float act_d2(int act, float z, float beta, float dh) {
    if (act == AM_ACT_SOFTPLUS) return z * beta > 20.0f ? 0.0f : beta * dh * (1.0f - dh);
    if (act == AM_ACT_SIGMOID) {
        const float y = 1.0f / (1.0f + expf(-z));
        return dh * (1.0f - 2.0f * y);
    }
    return 0.0f;
}

// Activation and its derivative at z together (one exp for softplus).
__device__ __forceinline__ void act_fb(int act, float z, float beta, float& h, float& dh) {
    if (act == AM_ACT_SOFTPLUS) {
        softplus_fb(z, beta, h, dh);
    } else {
        h = act_fwd(act, z, beta);
        dh = act_bwd(act, z, beta);
    }
}

__device__ __forceinline__ void act_fb4(int act, v4f z, float beta, v4f& h, v4f& dh) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        float a, d;
        act_fb(act, z[r], beta, a, d);
        h[r] = a;
        dh[r] = d;
    }
}

__device__ __forceinline__ v4f act4(int act, v4f z, float beta) {
    v4f h;
#pragma unroll
    for (int r = 0; r < 4; ++r) h[r] = act_fwd(act, z[r], beta);
    return h;
}

// Bias into a D tile: lane group k, register r -> row 16t + 4k + r.
__device__ __forceinline__ v4f bias_tile(const float* __restrict__ b, int t, int nrows) {
    const int k = lane_id() >> 4;
    v4f v;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = 16 * t + 4 * k + r;
        v[r] = row < nrows ? b[row] : 0.0f;
    }
    return v;
}

// acc[ot] += W[16ot + i][krow(s,k)] * b[s]
template <int KS, int OT, class Frag, int PR = kF32>
__device__ __forceinline__ void mm_fwd(const float* __restrict__ W, int ld, Frag b, v4f* acc) {
    const int l = lane_id();
    const int i = l & 15, k = l >> 4;
    if constexpr (PR == kF32) {
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const float bs = b[s];
#pragma unroll
            for (int ot = 0; ot < OT; ++ot) {
                const float a = W[(16 * ot + i) * ld + krow(s, k)];
                acc[ot] = mfma4(a, bs, acc[ot]);
            }
        }
    } else {
#pragma unroll
        for (int u = 0; u < (KS + 3) / 4; ++u) {
            v4f bv;
#pragma unroll
            for (int q = 0; q < 4; ++q) bv[q] = 4 * u + q < KS ? b[4 * u + q] : 0.0f;
#pragma unroll
            for (int ot = 0; ot < OT; ++ot) {
                v4f av;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    av[q] = 4 * u + q < KS ? W[(16 * ot + i) * ld + krow(4 * u + q, k)] : 0.0f;
                acc[ot] = mfma_bf16_step(av, bv, acc[ot]);
            }
        }
    }
}

// gin[it] += W[krow(s,k)][16it + i] * g[s]   (W^T * G)
template <int KS, int IT, class Frag, int PR = kF32>
__device__ __forceinline__ void mm_bwd_in(const float* __restrict__ W, int ld, Frag g, v4f* gin) {
    const int l = lane_id();
    const int i = l & 15, k = l >> 4;
    if constexpr (PR == kF32) {
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const float gs = g[s];
#pragma unroll
            for (int it = 0; it < IT; ++it) {
                const float a = W[krow(s, k) * ld + 16 * it + i];
                gin[it] = mfma4(a, gs, gin[it]);
            }
        }
    } else {
#pragma unroll
        for (int u = 0; u < (KS + 3) / 4; ++u) {
            v4f gv;
#pragma unroll
            for (int q = 0; q < 4; ++q) gv[q] = 4 * u + q < KS ? g[4 * u + q] : 0.0f;
#pragma unroll
            for (int it = 0; it < IT; ++it) {
                v4f av;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    av[q] = 4 * u + q < KS ? W[krow(4 * u + q, k) * ld + 16 * it + i] : 0.0f;
                gin[it] = mfma_bf16_step(av, gv, gin[it]);
            }
        }
    }
}

// Row stride of the transpose scratch: 20 keeps both its row writes (lane
// group k row 4k + r) and its column reads (lane i row i, column 4q + k)
// free of bank conflicts (17 put four lanes on a bank when writing).
constexpr int kScratchLd = 20;

// dW[ot*IT + it] += G (OT D tiles, rows = out) x X^T (KS B fragments, rows =
// input row space) summed over the 16 points of the tile.  Transposes go
// through the wave-private LDS scratch (16*(OT+IT) rows x kScratchLd).
// Result layout: lane (i,k), register r -> dW[16ot + 4k + r][16it + i].
template <int OT, int IT, int KS, class Frag, int PR = kF32>
__device__ __forceinline__ void mm_dw(const v4f* G, Frag X, float* scratch, v4f* dW) {
    const int l = lane_id();
    const int i = l & 15, k = l >> 4;
    float* Gs = scratch;
    float* Xs = scratch + 16 * OT * kScratchLd;
#pragma unroll
    for (int t = 0; t < OT; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) Gs[(16 * t + 4 * k + r) * kScratchLd + i] = G[t][r];
#pragma unroll
    for (int s = 0; s < 4 * IT; ++s) Xs[krow(s, k) * kScratchLd + i] = s < KS ? X[s] : 0.0f;
    wave_sync();
#pragma unroll
    for (int ot = 0; ot < OT; ++ot) {
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            v4f acc = dW[ot * IT + it];
            if constexpr (PR == kF32) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float a = Gs[(16 * ot + i) * kScratchLd + 4 * q + k];
                    const float b = Xs[(16 * it + i) * kScratchLd + 4 * q + k];
                    acc = mfma4(a, b, acc);
                }
            } else {
                v4f av, bv;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    av[q] = Gs[(16 * ot + i) * kScratchLd + 4 * q + k];
                    bv[q] = Xs[(16 * it + i) * kScratchLd + 4 * q + k];
                }
                acc = mfma_bf16_step(av, bv, acc);
            }
            dW[ot * IT + it] = acc;
        }
    }
    wave_sync();
}

// Sum of a D tile's rows over the 16 points (bias gradient), all lanes of the
// group receive the sums of their 4 rows.
__device__ __forceinline__ v4f rowsum_tile(v4f g) {
    v4f s;
#pragma unroll
    for (int r = 0; r < 4; ++r) s[r] = sum_in_group(g[r]);
    return s;
}

}  // namespace mlp
}  // namespace am
