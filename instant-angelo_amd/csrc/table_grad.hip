// Deterministic table-gradient reduction: stable radix sort of (row, value)
// pairs, then a fixed-order segmented reduce (see table_grad.h).
#include <rocprim/device/device_radix_sort.hpp>
#include "table_grad.h"

namespace am {

int table_key_bits(uint32_t n_entries) {
    int b = 1;
    while (b < 32 && (1ull << b) <= (uint64_t)n_entries) ++b;
    return b;
}

// Row keys of 25-27 bits (tables of 16.7 M - 134 M rows) sort in three
// 9-bit onesweep passes instead of four 8-bit ones.
using Sort9Config = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<rocprim::kernel_config<1024, 8>,
                                        rocprim::kernel_config<1024, 8>, 9,
                                        rocprim::block_radix_rank_algorithm::match>>;

template <class VT>
static hipError_t sort_pairs(void* tmp, size_t& bytes, TableReduceWs& w, uint32_t** kout,
                             VT** vout, hipStream_t s) {
    rocprim::double_buffer<uint32_t> kb(w.keys[0], w.keys[1]);
    rocprim::double_buffer<VT> vb(reinterpret_cast<VT*>(w.vals[0]), reinterpret_cast<VT*>(w.vals[1]));
    hipError_t e;
    if (w.end_bit > 24 && w.end_bit <= 27)
        e = rocprim::radix_sort_pairs<Sort9Config>(tmp, bytes, kb, vb, (size_t)w.m, 0u,
                                                   (unsigned)w.end_bit, s, false);
    else
        e = rocprim::radix_sort_pairs(tmp, bytes, kb, vb, (size_t)w.m, 0u, (unsigned)w.end_bit, s,
                                      false);
    if (kout) *kout = kb.current();
    if (vout) *vout = vb.current();
    return e;
}

int table_reduce_carve(WsCarver& c, int F, int64_t m, uint32_t n_entries, TableReduceWs* out,
                       bool half_vals) {
    TableReduceWs w{};
    w.m = m;
    w.end_bit = table_key_bits(n_entries);
    w.half_vals = half_vals;
    const size_t vbytes = (size_t)F * (half_vals ? sizeof(__half) : sizeof(float));
    w.keys[0] = c.take<uint32_t>((size_t)m);
    w.keys[1] = c.take<uint32_t>((size_t)m);
    w.vals[0] = c.take<char>((size_t)m * vbytes);
    w.vals[1] = c.take<char>((size_t)m * vbytes);
    const int64_t nc = ceil_div(m, kSegChunk);
    w.lead = c.take<float>((size_t)nc * F);
    w.tail = c.take<float>((size_t)nc * F);
    w.lead_full = c.take<uint8_t>((size_t)nc);
    w.tail_open = c.take<uint8_t>((size_t)nc);
    size_t sb = 0;
    hipError_t e = hipSuccess;
    if (half_vals && F != 2) {
        set_error("fp16 gradient pairs need n_features = 2 (got %d)", F);
        return AM_ERR_UNSUPPORTED;
    }
    switch (half_vals ? -F : F) {
        case 1: e = sort_pairs<Vals<1>>(nullptr, sb, w, nullptr, nullptr, 0); break;
        case 2: e = sort_pairs<Vals<2>>(nullptr, sb, w, nullptr, nullptr, 0); break;
        case 4: e = sort_pairs<Vals<4>>(nullptr, sb, w, nullptr, nullptr, 0); break;
        case 8: e = sort_pairs<Vals<8>>(nullptr, sb, w, nullptr, nullptr, 0); break;
        case -2: e = sort_pairs<ValsH<2>>(nullptr, sb, w, nullptr, nullptr, 0); break;
        default: set_error("unsupported n_features %d", F); return AM_ERR_UNSUPPORTED;
    }
    if (e != hipSuccess) {
        set_error("radix sort size query failed: %s", hipGetErrorString(e));
        return AM_ERR_HIP;
    }
    w.sort_bytes = sb;
    w.sort_tmp = c.take<char>(sb + 256);
    *out = w;
    return AM_OK;
}

__global__ void k_fill_f32(float* __restrict__ p, int64_t n, float v) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// Pass A: each thread reduces kSegChunk consecutive sorted items.  Segments
// wholly inside the chunk are written; the piece continuing a segment from the
// previous chunk goes to lead[t] (lead_full: the chunk is entirely that piece
// and the segment goes on); a segment starting here and continuing goes to
// tail[t] (tail_open).  The block first stages its 256 chunks through LDS
// with coalesced loads (kSegChunk is odd, so per-thread LDS strides spread
// over the banks).
template <int F>
struct SegCfg {
    static constexpr int threads = F <= 2 ? 256 : 64;
};

template <int F, class VT>
__global__ void __launch_bounds__(SegCfg<F>::threads) k_seg_a(
    const uint32_t* __restrict__ keys, const VT* __restrict__ vals, int64_t m,
    uint32_t n_entries, float* __restrict__ out, float* __restrict__ lead,
    float* __restrict__ tail, uint8_t* __restrict__ lead_full, uint8_t* __restrict__ tail_open) {
    constexpr int kSegThreads = SegCfg<F>::threads;
    __shared__ uint32_t sk[kSegThreads * kSegChunk + 1];
    __shared__ float sv[kSegThreads * kSegChunk * F];
    const int64_t blk0 = (int64_t)blockIdx.x * kSegThreads * kSegChunk;
    const int64_t blk_n = m - blk0 < (int64_t)kSegThreads * kSegChunk ? m - blk0
                                                                     : (int64_t)kSegThreads * kSegChunk;
    for (int q = threadIdx.x; q < blk_n; q += kSegThreads) sk[q] = keys[blk0 + q];
    if constexpr (sizeof(VT) == F * sizeof(float)) {
        const float* vf = reinterpret_cast<const float*>(vals) + blk0 * F;
        for (int q = threadIdx.x; q < blk_n * F; q += kSegThreads) sv[q] = vf[q];
    } else {
        const __half* vh = reinterpret_cast<const __half*>(vals) + blk0 * F;
        for (int q = threadIdx.x; q < blk_n * F; q += kSegThreads) sv[q] = __half2float(vh[q]);
    }
    __syncthreads();
    const int64_t t = (int64_t)blockIdx.x * kSegThreads + threadIdx.x;
    const int64_t nc = (m + kSegChunk - 1) / kSegChunk;
    if (t >= nc) return;
    const int lb = threadIdx.x * kSegChunk;                 // local begin
    const int64_t b = t * kSegChunk;
    const int64_t e = b + kSegChunk < m ? b + kSegChunk : m;
    const int le = lb + (int)(e - b);
    bool is_lead = b > 0 && sk[lb] == (lb > 0 ? sk[lb - 1] : keys[b - 1]);
    uint32_t cur = sk[lb];
    float acc[F];
#pragma unroll
    for (int f = 0; f < F; ++f) acc[f] = 0.0f;
    for (int j = lb; j < le; ++j) {
        const uint32_t k = sk[j];
        if (k != cur) {
            if (is_lead) {
#pragma unroll
                for (int f = 0; f < F; ++f) lead[t * F + f] = acc[f];
                is_lead = false;
            } else if (cur < n_entries) {
#pragma unroll
                for (int f = 0; f < F; ++f) out[(int64_t)cur * F + f] = acc[f];
            }
            cur = k;
#pragma unroll
            for (int f = 0; f < F; ++f) acc[f] = 0.0f;
        }
#pragma unroll
        for (int f = 0; f < F; ++f) acc[f] += sv[j * F + f];
    }
    const bool cont_out = e < m && keys[e] == cur;
    uint8_t lf = 0, to = 0;
    if (is_lead) {
#pragma unroll
        for (int f = 0; f < F; ++f) lead[t * F + f] = acc[f];
        lf = cont_out ? 1 : 0;
    } else if (cont_out) {
#pragma unroll
        for (int f = 0; f < F; ++f) tail[t * F + f] = acc[f];
        to = 1;
    } else if (cur < n_entries) {
#pragma unroll
        for (int f = 0; f < F; ++f) out[(int64_t)cur * F + f] = acc[f];
    }
    lead_full[t] = lf;
    tail_open[t] = to;
}

// Pass B: a segment spanning chunks is summed by the chunk it starts in:
// its tail piece, then each following chunk's lead piece, in chunk order.
template <int F>
__global__ void k_seg_b(const uint32_t* __restrict__ keys, int64_t m, uint32_t n_entries,
                        float* __restrict__ out, const float* __restrict__ lead,
                        const float* __restrict__ tail, const uint8_t* __restrict__ lead_full,
                        const uint8_t* __restrict__ tail_open) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nc = (m + kSegChunk - 1) / kSegChunk;
    if (t >= nc || !tail_open[t]) return;
    const int64_t last = (t + 1) * kSegChunk - 1;
    const uint32_t key = keys[last];
    if (key >= n_entries) return;
    float acc[F];
#pragma unroll
    for (int f = 0; f < F; ++f) acc[f] = tail[t * F + f];
    for (int64_t u = t + 1; u < nc; ++u) {
#pragma unroll
        for (int f = 0; f < F; ++f) acc[f] += lead[u * F + f];
        if (!lead_full[u]) break;
    }
#pragma unroll
    for (int f = 0; f < F; ++f) out[(int64_t)key * F + f] = acc[f];
}

int fill_f32(float* p, int64_t n, float v, hipStream_t s) {
    if (n <= 0) return AM_OK;
    int blocks = (int)(ceil_div(n, 256) < 8192 ? ceil_div(n, 256) : 8192);
    hipLaunchKernelGGL(k_fill_f32, dim3(blocks), dim3(256), 0, s, p, n, v);
    AM_LAUNCH_CHECK();
    return AM_OK;
}

template <int F, class VT>
static int run_f(TableReduceWs& w, uint32_t n_entries, float* dtable, hipStream_t s) {
    if (int rc = fill_f32(dtable, (int64_t)n_entries * F, 0.0f, s)) return rc;
    if (w.m == 0) return AM_OK;
    uint32_t* ks = nullptr;
    VT* vs = nullptr;
    size_t sb = w.sort_bytes;
    hipError_t e;
    {
        KernelTimer kt_("radix_sort", s);
        e = sort_pairs<VT>(w.sort_tmp, sb, w, &ks, &vs, s);
    }
    if (e != hipSuccess) {
        set_error("radix sort failed: %s", hipGetErrorString(e));
        return AM_ERR_HIP;
    }
    const int64_t nc = ceil_div(w.m, kSegChunk);
    const int gb = (int)ceil_div(nc, 256);
    {
        KernelTimer kt_("seg_reduce", s);
        constexpr int T = SegCfg<F>::threads;
        hipLaunchKernelGGL((k_seg_a<F, VT>), dim3((unsigned)ceil_div(nc, T)), dim3(T), 0, s, ks, vs, w.m,
                           n_entries, dtable, w.lead, w.tail, w.lead_full, w.tail_open);
    }
    AM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_seg_b<F>, dim3(gb), dim3(256), 0, s, ks, w.m, n_entries, dtable, w.lead,
                       w.tail, w.lead_full, w.tail_open);
    AM_LAUNCH_CHECK();
    return AM_OK;
}

int table_reduce_run(TableReduceWs& w, int F, uint32_t n_entries, float* dtable,
                     hipStream_t stream) {
    if (w.half_vals) {
        if (F == 2) return run_f<2, ValsH<2>>(w, n_entries, dtable, stream);
        set_error("fp16 gradient pairs need n_features = 2 (got %d)", F);
        return AM_ERR_UNSUPPORTED;
    }
    switch (F) {
        case 1: return run_f<1, Vals<1>>(w, n_entries, dtable, stream);
        case 2: return run_f<2, Vals<2>>(w, n_entries, dtable, stream);
        case 4: return run_f<4, Vals<4>>(w, n_entries, dtable, stream);
        case 8: return run_f<8, Vals<8>>(w, n_entries, dtable, stream);
        default: set_error("unsupported n_features %d", F); return AM_ERR_UNSUPPORTED;
    }
}

}  // namespace am
