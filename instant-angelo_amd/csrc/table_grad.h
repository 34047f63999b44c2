// Deterministic table-gradient reduction: (row key, value) pairs emitted by a
// backward kernel -> stable radix sort by key -> segmented reduce in a fixed
// order -> one write per touched row; untouched rows are written 0.
//
// Two-feature tables (the hash grids here) sort only the key bits above
// `begin_bit` (two passes for 25-27-bit keys instead of three): the
// pairs of each bucket of 2^begin_bit rows are then contiguous and in
// emission order, and one wave per bucket finishes the job with an in-wave
// stable split sort and a segmented scan into per-row accumulators in LDS.
//
// The emitters write pairs into carve.keys[0] / carve.vals[0].  Keys equal to
// sentinel() mark empty slots; they sort last and are skipped.
#pragma once
#include <hip/hip_fp16.h>
#include "am_common.h"

namespace am {

template <int F>
struct Vals {
    float v[F];
};

// Pair values stored in fp16 (the mixed-precision form: loss-scaled
// gradients, summed in fp32 by the segmented reduce).
template <int F>
struct ValsH {
    __half v[F];
};

constexpr int kSegChunk = 17;  // items reduced sequentially per thread (odd: LDS banks)

struct TableReduceWs {
    uint32_t* keys[2];
    void* vals[2];
    void* sort_tmp;
    size_t sort_bytes;
    float* lead;
    float* tail;
    uint8_t* lead_full;
    uint8_t* tail_open;
    int64_t* bucket_start;  // [n_buckets + 1] (bucketed form)
    uint64_t* units;        // [n_buckets + 1] packed work-unit counts
    uint64_t* unit_off;     // [n_buckets + 1] their exclusive scan
    float* partial;         // partial sums of multi-unit buckets
    void* scan_tmp;
    size_t scan_bytes;
    int64_t n_buckets;      // 0: full sort + chunked segmented reduce
    int64_t m;
    int begin_bit;          // key bits sorted: [begin_bit, end_bit)
    int end_bit;
    bool half_vals;  // values are ValsH<F> instead of Vals<F>
    // The emitter wrote the pairs partitioned by key bits [9, 17) already
    // (in emission order inside each partition): the sort makes one pass over
    // bits [17, end_bit), or none for end_bit <= 17.
    bool prepart;
};

// First key bit the sort of prepartitioned pairs orders (bits [9, 17) are
// the emitter's partition digit).
constexpr int kPrepartBits = 17;

// Bucket bits of the bucketed form: the low key bits left unsorted (at most
// 9, so a wave's accumulators span at most 512 rows).
int table_bucket_bits(int end_bit);

// Number of key bits: smallest b with 2^b > n_entries (so the sentinel
// (2^b - 1) is >= n_entries and never a valid row).
int table_key_bits(uint32_t n_entries);
static inline uint32_t table_sentinel(int end_bit) { return (uint32_t)((1ull << end_bit) - 1ull); }

// Carves the reduction workspace for m pairs of F features out of `c`.
// Returns the sort's temp byte count through TableReduceWs (0 when c has no
// base pointer: size query; the sort size query needs a device).
int table_reduce_carve(WsCarver& c, int F, int64_t m, uint32_t n_entries, TableReduceWs* out,
                       bool half_vals = false, bool prepart = false);

// Writes v to p[0..n) on `stream` (a fill kernel, safe under graph capture).
int fill_f32(float* p, int64_t n, float v, hipStream_t stream);

// Runs sort + segmented reduce; dtable [n_entries, F] fully overwritten.
int table_reduce_run(TableReduceWs& w, int F, uint32_t n_entries, float* dtable,
                     hipStream_t stream);

}  // namespace am
