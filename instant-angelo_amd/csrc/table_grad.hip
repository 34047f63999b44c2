// Deterministic table-gradient reduction: stable radix sort of (row, value)
// pairs, then a fixed-order segmented reduce (see table_grad.h).
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include "table_grad.h"

namespace am {

constexpr int64_t kUnitItems = 16384;   // sorted items per work unit (256 chunks)

// Partial-sum slots the units of multi-unit buckets can need: each such
// bucket holds more than kUnitItems items and fewer than 2n / kUnitItems units.
static int64_t bucket_slots(int64_t m) { return 2 * (m / kUnitItems) + 1; }

int table_key_bits(uint32_t n_entries) {
    int b = 1;
    while (b < 32 && (1ull << b) <= (uint64_t)n_entries) ++b;
    return b;
}

// Bucketed form: up to 9 low key bits stay unsorted.  Keys of up to 25 bits
// sort their high bits [9, end) in one or two 8-bit onesweep passes (the
// fastest form measured for 16 bits: 8-bit digits rank with less LDS than
// 9-bit ones); 26-27-bit keys in two 9-bit passes, longer ones in three.
int table_bucket_bits(int end_bit) {
    if (end_bit <= 9) return 0;
    if (end_bit <= 25) return 9;
    const int passes = (end_bit - 9 + 8) / 9;
    return end_bit - 9 * passes;
}

static int digit_bits(const TableReduceWs& w) {
    return (w.n_buckets > 0 && w.begin_bit == 9) ? 8 : 9;
}

// Onesweep configurations: 1024-thread blocks of 8 items per thread for the
// sort passes, 32 items per thread for the histogram pass.
using Sort9Config = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<rocprim::kernel_config<1024, 32>,
                                        rocprim::kernel_config<1024, 8>, 9,
                                        rocprim::block_radix_rank_algorithm::match>>;
using Sort8Config = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<rocprim::kernel_config<1024, 32>,
                                        rocprim::kernel_config<1024, 8>, 8,
                                        rocprim::block_radix_rank_algorithm::match>>;

template <class VT>
static hipError_t sort_pairs(void* tmp, size_t& bytes, TableReduceWs& w, uint32_t** kout,
                             VT** vout, hipStream_t s) {
    rocprim::double_buffer<uint32_t> kb(w.keys[0], w.keys[1]);
    rocprim::double_buffer<VT> vb(reinterpret_cast<VT*>(w.vals[0]), reinterpret_cast<VT*>(w.vals[1]));
    hipError_t e;
    if (w.prepart) {   // one 8-bit pass over [17, end_bit), or none
        e = hipSuccess;
        if (w.end_bit > kPrepartBits)
            e = rocprim::radix_sort_pairs<Sort8Config>(tmp, bytes, kb, vb, (size_t)w.m,
                                                       (unsigned)kPrepartBits,
                                                       (unsigned)w.end_bit, s, false);
        else if (tmp == nullptr)
            bytes = 0;
    } else if (w.n_buckets > 0 && digit_bits(w) == 8)
        e = rocprim::radix_sort_pairs<Sort8Config>(tmp, bytes, kb, vb, (size_t)w.m,
                                                   (unsigned)w.begin_bit, (unsigned)w.end_bit, s,
                                                   false);
    else if (w.n_buckets > 0 || (w.end_bit > 24 && w.end_bit <= 27))
        e = rocprim::radix_sort_pairs<Sort9Config>(tmp, bytes, kb, vb, (size_t)w.m,
                                                   (unsigned)w.begin_bit, (unsigned)w.end_bit, s,
                                                   false);
    else
        e = rocprim::radix_sort_pairs(tmp, bytes, kb, vb, (size_t)w.m, 0u, (unsigned)w.end_bit, s,
                                      false);
    if (kout) *kout = kb.current();
    if (vout) *vout = vb.current();
    return e;
}

int table_reduce_carve(WsCarver& c, int F, int64_t m, uint32_t n_entries, TableReduceWs* out,
                       bool half_vals, bool prepart) {
    TableReduceWs w{};
    w.m = m;
    w.end_bit = table_key_bits(n_entries);
    w.half_vals = half_vals;
    if (prepart && (F != 2 || table_bucket_bits(w.end_bit) != 9 || w.end_bit > 25)) {
        set_error("prepartitioned pairs need two features and 9 < key bits <= 25 (got %d, %d)",
                  F, w.end_bit);
        return AM_ERR_INTERNAL;
    }
    w.prepart = prepart;
    const size_t vbytes = (size_t)F * (half_vals ? sizeof(__half) : sizeof(float));
    w.keys[0] = c.take<uint32_t>((size_t)m);
    w.keys[1] = c.take<uint32_t>((size_t)m);
    w.vals[0] = c.take<char>((size_t)m * vbytes);
    w.vals[1] = c.take<char>((size_t)m * vbytes);
    if (F == 2) {   // bucketed form
        w.begin_bit = table_bucket_bits(w.end_bit);
        w.n_buckets = ceil_div((int64_t)n_entries, (int64_t)1 << w.begin_bit);
        w.bucket_start = c.take<int64_t>((size_t)w.n_buckets + 1);
        w.units = c.take<uint64_t>((size_t)w.n_buckets + 1);
        w.unit_off = c.take<uint64_t>((size_t)w.n_buckets + 1);
        w.partial = c.take<float>((size_t)bucket_slots(m) * (2u << w.begin_bit));
        size_t scb = 0;
        if (rocprim::exclusive_scan(nullptr, scb, (uint64_t*)nullptr, (uint64_t*)nullptr,
                                    (uint64_t)0, (size_t)w.n_buckets + 1,
                                    rocprim::plus<uint64_t>(), 0, false) != hipSuccess) {
            set_error("bucket unit scan size query failed");
            return AM_ERR_HIP;
        }
        w.scan_bytes = scb;
        w.scan_tmp = c.take<char>(scb + 256);
    } else {        // chunked segmented reduce
        const int64_t nc = ceil_div(m, kSegChunk);
        w.lead = c.take<float>((size_t)nc * F);
        w.tail = c.take<float>((size_t)nc * F);
        w.lead_full = c.take<uint8_t>((size_t)nc);
        w.tail_open = c.take<uint8_t>((size_t)nc);
    }
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

// Bucketed form.  start[k] = first sorted item of bucket k (keys sorted on
// bits >= B, so key >> B is non-decreasing), k in [0, n_buckets].
__global__ void k_bucket_bounds(const uint32_t* __restrict__ keys, int64_t m, int B,
                                int64_t n_buckets, int64_t* __restrict__ start) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t > n_buckets) return;
    // start[n_buckets] is searched too: sentinel keys of empty slots may
    // lie past the last bucket and must stay out of its range
    int64_t lo = 0, hi = m;
    while (lo < hi) {
        const int64_t mid = lo + ((hi - lo) >> 1);
        if ((int64_t)(keys[mid] >> B) < t) lo = mid + 1;
        else hi = mid;
    }
    start[t] = lo;
}

__device__ __forceinline__ int popc_below(uint64_t mask) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                          __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// One round of a stable split of the wave's 64 keys on bit 6 + `bit`: lanes
// with the bit clear keep their order at the front (destination = clear
// lanes below), set ones follow (all clear lanes + set lanes below).
// Branch-free: both destinations come from one count of set lanes below.
__device__ __forceinline__ uint32_t split_round(uint32_t rk, int bit, int lane) {
    const bool one = (rk >> (6 + bit)) & 1u;
    const uint64_t ones = __ballot(one);
    const int below = popc_below(ones);
    const int dst = one ? (64 - (int)__popcll(ones)) + below : lane - below;
    return (uint32_t)__builtin_amdgcn_ds_permute(dst * 4, (int)rk);
}

__device__ __forceinline__ float permute_f(int dst_byte, float v) {
    return __int_as_float(__builtin_amdgcn_ds_permute(dst_byte, __float_as_int(v)));
}

__device__ __forceinline__ float bpermute_f(int src_byte, float v) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src_byte, __float_as_int(v)));
}

__device__ __forceinline__ void load_pair(const Vals<2>* v, int64_t i, float& a, float& b) {
    const float2 x = *reinterpret_cast<const float2*>(v + i);
    a = x.x;
    b = x.y;
}

__device__ __forceinline__ void load_pair(const ValsH<2>* v, int64_t i, float& a, float& b) {
    const __half2 x = *reinterpret_cast<const __half2*>(v + i);
    a = __low2float(x);
    b = __high2float(x);
}

// One step of a segmented inclusive scan over lanes sorted by r: the DPP
// source lane's partial is added when it holds the same row (lanes whose
// source is outside the row, or whose row the mask excludes, read the
// never-matching sentinel).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ void seg_step(uint32_t r, float& a, float& b) {
    const uint32_t rs = (uint32_t)__builtin_amdgcn_update_dpp((int)0xffffffffu, (int)r, CTRL,
                                                              ROW_MASK, 0xf, false);
    const float as = __int_as_float(
        __builtin_amdgcn_update_dpp(0, __float_as_int(a), CTRL, ROW_MASK, 0xf, false));
    const float bs = __int_as_float(
        __builtin_amdgcn_update_dpp(0, __float_as_int(b), CTRL, ROW_MASK, 0xf, false));
    if (rs == r) {
        a += as;
        b += bs;
    }
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kBucketWaves = 4;
constexpr int kMaxBucketBits = 9;

// Work units: a bucket of n items is cut into max(1, ceil(n / kUnitItems))
// consecutive ranges, so a bucket of a coarse (dense, heavily shared) level
// is spread over many waves.  units[b] packs the bucket's unit count (low 32
// bits) and, for a bucket of several units, the same count again (high 32
// bits: its partial-sum slots); an exclusive scan of the packed words gives
// each bucket's first unit and first slot at once.  units[n_buckets] = 0, so
// the scan's last entry holds the totals.
__global__ void k_bucket_units(const int64_t* __restrict__ start, int64_t n_buckets,
                               uint64_t* __restrict__ units) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t > n_buckets) return;
    uint64_t v = 0;
    if (t < n_buckets) {
        const int64_t n = start[t + 1] - start[t];
        const uint64_t ns = n <= kUnitItems ? 1u : (uint64_t)((n + kUnitItems - 1) / kUnitItems);
        v = ns | (ns > 1 ? ns << 32 : 0);
    }
    units[t] = v;
}

// Reduces sorted items [s, e) of one bucket into the wave's LDS accumulators
// acc[2 << B].  Each chunk of 64 items (in emission order) is stable-sorted
// in registers by its row within the bucket (B split rounds of ballot +
// permute; a partial chunk sorts one more bit that moves its empty lanes
// last), reduced by a segmented inclusive scan (DPP row shifts and
// broadcasts: a fixed tree per segment), and each segment's total is added
// to the row's accumulator by the segment's last lane (rows are unique per
// chunk, chunks go in order): the summation order is a fixed function of the
// sorted input.
template <class VT>
__device__ __forceinline__ void reduce_range(const uint32_t* __restrict__ keys,
                                             const VT* __restrict__ vals, int64_t s, int64_t e,
                                             int B, float* acc) {
    const int lane = threadIdx.x & 63;
    const int rows = 1 << B;
    uint32_t nk = 0;
    float na = 0.0f, nb = 0.0f;
    if (s + lane < e) {
        nk = keys[s + lane];
        load_pair(vals, s + lane, na, nb);
    }
    for (int64_t c = s; c < e; c += 64) {
        const bool valid = c + lane < e;
        const uint32_t r0 = valid ? (nk & (uint32_t)(rows - 1)) : (uint32_t)rows;
        const float a0 = valid ? na : 0.0f, b0 = valid ? nb : 0.0f;
        if (c + 64 + lane < e) {   // prefetch the next chunk
            nk = keys[c + 64 + lane];
            load_pair(vals, c + 64 + lane, na, nb);
        }
        // stable split sort of (row << 6 | source lane): one permute per
        // round; the values follow once at the end.  A partial chunk sorts
        // one more bit (its idle lanes carry row `rows`).
        uint32_t rk = (r0 << 6) | (uint32_t)lane;
        if (B == 9) {   // the bucketed sort's digit: unrolled
#pragma unroll
            for (int bit = 0; bit < 9; ++bit) rk = split_round(rk, bit, lane);
            if (e - c < 64) rk = split_round(rk, 9, lane);
        } else {
            const int nbits = (e - c < 64) ? B + 1 : B;
            for (int bit = 0; bit < nbits; ++bit) rk = split_round(rk, bit, lane);
        }
        const uint32_t r = rk >> 6;
        const int from = (int)(rk & 63u) * 4;
        float a = bpermute_f(from, a0);
        float b = bpermute_f(from, b0);
        // segmented inclusive scan over the sorted lanes: row shifts of
        // 1, 2, 4, 8 within 16-lane rows, then the last lane of row 0
        // (row 2) into row 1 (row 3), then lane 31 into rows 2 and 3
        seg_step<0x111, 0xf>(r, a, b);
        seg_step<0x112, 0xf>(r, a, b);
        seg_step<0x114, 0xf>(r, a, b);
        seg_step<0x118, 0xf>(r, a, b);
        seg_step<0x142, 0xa>(r, a, b);
        seg_step<0x143, 0xc>(r, a, b);
        const uint32_t rn = (uint32_t)__builtin_amdgcn_ds_bpermute(((lane + 1) & 63) * 4, (int)r);
        if ((lane == 63 || rn != r) && r < (uint32_t)rows) {
            acc[2 * r] += a;
            acc[2 * r + 1] += b;
        }
        wave_lds_sync();
    }
}

// One wave per work unit (grid-stride).  A single-unit bucket writes its
// rows to dtable, zeros included (dtable needs no fill); a unit of a
// multi-unit bucket writes its accumulators to its partial slot.
template <class VT>
__global__ void __launch_bounds__(64 * kBucketWaves) k_bucket_reduce(
    const uint32_t* __restrict__ keys, const VT* __restrict__ vals,
    const int64_t* __restrict__ start, const uint64_t* __restrict__ uoff, int64_t n_buckets,
    int B, uint32_t n_entries, float* __restrict__ out, float* __restrict__ partial) {
    __shared__ float acc_s[kBucketWaves][2 << kMaxBucketBits];
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    float* acc = acc_s[wave];
    const int rows = 1 << B;
    const int64_t total = (int64_t)(uint32_t)uoff[n_buckets];
    for (int64_t u = (int64_t)blockIdx.x * kBucketWaves + wave; u < total;
         u += (int64_t)gridDim.x * kBucketWaves) {
        // the bucket: the last b with first unit <= u
        int64_t lo = 0, hi = n_buckets - 1;
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) >> 1;
            if ((int64_t)(uint32_t)uoff[mid] <= u) lo = mid;
            else hi = mid - 1;
        }
        const int64_t bk = lo;
        const uint64_t ub = uoff[bk];
        const int64_t sub = u - (int64_t)(uint32_t)ub;
        const int64_t ns = (int64_t)(uint32_t)uoff[bk + 1] - (int64_t)(uint32_t)ub;
        const int64_t s = start[bk] + sub * kUnitItems;
        const int64_t e = s + kUnitItems < start[bk + 1] ? s + kUnitItems : start[bk + 1];
        for (int q = lane; q < 2 * rows; q += 64) acc[q] = 0.0f;
        wave_lds_sync();
        reduce_range(keys, vals, s, e, B, acc);
        if (ns == 1) {
            const int64_t row0 = bk << B;
            for (int q = lane; q < rows; q += 64) {
                const int64_t row = row0 + q;
                if (row < (int64_t)n_entries)
                    *reinterpret_cast<float2*>(out + 2 * row) =
                        make_float2(acc[2 * q], acc[2 * q + 1]);
            }
        } else {
            float* dst = partial + ((int64_t)(ub >> 32) + sub) * (2 * rows);
            for (int q = lane; q < 2 * rows; q += 64) dst[q] = acc[q];
        }
        wave_lds_sync();
    }
}

// Multi-unit buckets: each row is the sum of its units' partials in unit
// order (one wave per bucket, grid-stride).
__global__ void __launch_bounds__(256) k_bucket_combine(const uint64_t* __restrict__ uoff,
                                                        int64_t n_buckets, int B,
                                                        uint32_t n_entries, float* __restrict__ out,
                                                        const float* __restrict__ partial) {
    const int lane = threadIdx.x & 63;
    const int rows = 1 << B;
    for (int64_t bk = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
         bk < n_buckets; bk += (int64_t)gridDim.x * (blockDim.x >> 6)) {
        const uint64_t ub = uoff[bk];
        const int64_t ns = (int64_t)(uint32_t)uoff[bk + 1] - (int64_t)(uint32_t)ub;
        if (ns < 2) continue;
        const float* src = partial + (int64_t)(ub >> 32) * (2 * rows);
        for (int q = lane; q < rows; q += 64) {
            const int64_t row = (bk << B) + q;
            if (row >= (int64_t)n_entries) continue;
            float a = 0.0f, b = 0.0f;
            for (int64_t j = 0; j < ns; ++j) {
                a += src[j * 2 * rows + 2 * q];
                b += src[j * 2 * rows + 2 * q + 1];
            }
            *reinterpret_cast<float2*>(out + 2 * row) = make_float2(a, b);
        }
    }
}

template <class VT>
static int run_bucketed(TableReduceWs& w, uint32_t n_entries, float* dtable, hipStream_t s) {
    uint32_t* ks = w.keys[0];
    VT* vs = reinterpret_cast<VT*>(w.vals[0]);
    if (w.m > 0) {
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
    }
    KernelTimer kt_("seg_reduce", s);
    const unsigned gb = (unsigned)ceil_div(w.n_buckets + 1, 256);
    hipLaunchKernelGGL(k_bucket_bounds, dim3(gb), dim3(256), 0, s, ks, w.m, w.begin_bit,
                       w.n_buckets, w.bucket_start);
    AM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_bucket_units, dim3(gb), dim3(256), 0, s, w.bucket_start, w.n_buckets,
                       w.units);
    AM_LAUNCH_CHECK();
    size_t scb = w.scan_bytes;
    hipError_t e = rocprim::exclusive_scan(w.scan_tmp, scb, w.units, w.unit_off, (uint64_t)0,
                                           (size_t)w.n_buckets + 1, rocprim::plus<uint64_t>(), s,
                                           false);
    if (e != hipSuccess) {
        set_error("bucket unit scan failed: %s", hipGetErrorString(e));
        return AM_ERR_HIP;
    }
    // units <= n_buckets + m / kUnitItems + 1; the kernel reads the total
    const int64_t max_units = w.n_buckets + w.m / kUnitItems + 1;
    const int64_t want = ceil_div(max_units, kBucketWaves);
    const unsigned grid = (unsigned)(want < 8192 ? want : 8192);
    hipLaunchKernelGGL((k_bucket_reduce<VT>), dim3(grid), dim3(64 * kBucketWaves), 0, s, ks, vs,
                       w.bucket_start, w.unit_off, w.n_buckets, w.begin_bit, n_entries, dtable,
                       w.partial);
    AM_LAUNCH_CHECK();
    const int64_t cw = ceil_div(w.n_buckets, 4);
    hipLaunchKernelGGL(k_bucket_combine, dim3((unsigned)(cw < 4096 ? cw : 4096)), dim3(256), 0, s,
                       w.unit_off, w.n_buckets, w.begin_bit, n_entries, dtable, w.partial);
    AM_LAUNCH_CHECK();
    return AM_OK;
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
    if (F == 2 && w.n_buckets > 0) {
        if (w.begin_bit > kMaxBucketBits) {
            set_error("bucket bits %d above %d", w.begin_bit, kMaxBucketBits);
            return AM_ERR_INTERNAL;
        }
        return w.half_vals ? run_bucketed<ValsH<2>>(w, n_entries, dtable, stream)
                           : run_bucketed<Vals<2>>(w, n_entries, dtable, stream);
    }
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
