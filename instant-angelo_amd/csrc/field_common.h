// Helpers shared by the fused geometry and radiance kernels.
#pragma once
#include "am_common.h"

namespace am {

static constexpr int kFieldBlocksMax = 1024;

static inline int field_blocks(int64_t n) {
    return persistent_blocks((n + 15) / 16, 4, kFieldBlocksMax);
}

// Grids of the SDF forward and of the pair passes (count, emit); their
// tiles' results do not depend on the grid.  Each is a whole number of
// rounds of the blocks the device holds at once on 256 CUs (the register
// counts of the fp32 16-level kernels): the forward two rounds of 512 (two
// waves a SIMD), the emission two rounds of 768 (three), the count one
// round of 2048 (eight); a partial last round measured slower.
static constexpr int kFwdBlocksMax = 1024;
static constexpr int kPairBlocksMax = 1536;
static constexpr int kCountBlocksMax = 2048;
static inline int fwd_blocks(int64_t n) { return persistent_blocks((n + 15) / 16, 4, kFwdBlocksMax); }
static inline int pair_blocks(int64_t n) { return persistent_blocks((n + 15) / 16, 4, kPairBlocksMax); }
// The backward passes of the SDF and radiance MLPs (over 200 VGPRs: two
// waves a SIMD, 512 resident blocks) in one round; the grid fixes the
// per-wave parameter-gradient rows and so their summation order, the same
// for every form of a kernel.
static constexpr int kBwdBlocksMax = 512;
static inline int bwd_blocks(int64_t n) { return persistent_blocks((n + 15) / 16, 4, kBwdBlocksMax); }
static inline int count_blocks(int64_t n) {
    return persistent_blocks((n + 15) / 16, 4, kCountBlocksMax);
}

// Widest output of the fused SDF MLP (five 16-row output tiles: the SDF and
// up to 79 features, e.g. the 65 outputs of the Neuralangelo configurations).
constexpr int kGeoMaxOut = 80;

static inline int check_mlp_common(const am_mlp_desc* m, int max_out = 16) {
    AM_CHECK_ARG(m != nullptr, "mlp descriptor is NULL");
    AM_CHECK_ARG(m->n_neurons == 64, "n_neurons %d: only 64 is implemented", m->n_neurons);
    AM_CHECK_ARG(m->n_hidden_layers == 1 || m->n_hidden_layers == 2,
                 "fused field kernels take 1 or 2 hidden layers (got %d)", m->n_hidden_layers);
    AM_CHECK_ARG(m->hidden_act == AM_ACT_RELU || m->hidden_act == AM_ACT_SOFTPLUS ||
                     m->hidden_act == AM_ACT_NONE,
                 "hidden_act %d", m->hidden_act);
    AM_CHECK_ARG(m->output_act == AM_ACT_NONE, "fused field kernels need output_act none");
    AM_CHECK_ARG(m->n_out >= 1 && m->n_out <= max_out, "n_out %d out of [1,%d]", m->n_out,
                 max_out);
    AM_CHECK_ARG(m->precision == AM_PREC_FP32 || m->precision == AM_PREC_MIXED, "precision %d",
                 m->precision);
    return AM_OK;
}

// Flat VanillaMLP parameter layout [W1,b1,(W2,b2),Wo,bo].
static inline void layout_params(int n_in, int n_out, int nh, int64_t* pw1, int64_t* pb1,
                                 int64_t* pw2, int64_t* pb2, int64_t* pwo, int64_t* pbo,
                                 int64_t* n_params) {
    int64_t o = 0;
    *pw1 = o; o += 64LL * n_in;
    *pb1 = o; o += 64;
    *pw2 = o; *pb2 = o;
    if (nh == 2) { *pw2 = o; o += 64 * 64; *pb2 = o; o += 64; }
    *pwo = o; o += 64LL * n_out;
    *pbo = o; o += n_out;
    *n_params = o;
}

}  // namespace am
