// Launchers of the descriptor-matching kernels (match.hip).
//
// The work is a batch of image pairs (a, b), each a brute-force comparison of
// two sets of 128-D f32 descriptors that stay resident in one array `desc`
// (row d = descriptor d, 128 floats).  Per pair the library finds, for every
// descriptor of a, its nearest and second-nearest neighbour in b (squared L2,
// summed in f32 in k order), the same for every descriptor of b against a,
// and keeps the mutual nearest neighbours that pass the symmetric ratio test
// (sfm-lite/src/matching.rs:22-98).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mt {

constexpr int TILE = 64;      // descriptors per tile side
constexpr int DIM = 128;      // descriptor length
constexpr uint32_t NONE = 0xffffffffu;

// best / second-best squared distance of one descriptor against a set, with
// the set index of the best; `b` = +inf and `i` = NONE when nothing is below
// +inf.
struct Best2 {
    float b;
    uint32_t i;
    float s;
    uint32_t pad;
};

// One pair of the batch.  Offsets are element offsets into the batch arrays.
struct PairDesc {
    int64_t a, b;          // first descriptor row of each set in `desc`
    int32_t na, nb;        // set sizes
    int32_t ta, tb;        // tiles along a and along b (ceil(n / TILE))
    int64_t tile0;         // first tile of this pair in the flattened tile space
    int64_t rpart, cpart;  // row partials [tb][na] and column partials [ta][nb]
    int64_t frow, fcol;    // per-descriptor results fwd[na], bwd[nb] (whole batch)
    int64_t lrow, lcol;    // the same rows / columns numbered within this launch's pair list
    int64_t out;           // first output slot (min(na, nb) slots per pair)
};

// Filtered search (bf16x3 MFMA products, exact re-evaluation of the few
// descriptors that can be nearest or second nearest).
constexpr int FTILE = 128;   // descriptors per filtered tile side
constexpr int CAP = 16;      // candidate slots per descriptor
// |exact f32 distance - MFMA estimate| <= EPS_REL * (|a|^2 + |b|^2) on
// descriptors whose nonzero components lie in [2^-40, 2^40] (safe sets)
constexpr float EPS_REL = 1.0f / 2048.0f;

// Per-descriptor squared norm and bf16 hi / lo split (x = hi + lo + r,
// |r| <= 2^-18 |x|) of rows [0, n); bad[set] |= 1 when a row of the set has
// a component outside {0} U [2^-40, 2^40] (or non-finite).
hipError_t launch_prep(const float* desc, int64_t n, const int64_t* set_off, int nsets, float* norm, uint16_t* hi,
                       uint16_t* lo, int* bad, hipStream_t st);

// Per row (or column) and tile of the filtered search: the two smallest
// upper bounds and a lower bound of every element's L in the tile; beside
// it a 128-bit mask of the tile's elements whose L is <= u2.
struct TileMeta {
    float u1, u2, lb;
    uint32_t pad;
};

// Buffers of one filtered launch (indices are launch-local rows / columns).
struct FiltBufs {
    TileMeta* rmeta;   // [tb][na] per row and column tile
    uint4* rmask;
    TileMeta* cmeta;   // [ta][nb] per column and row tile
    uint4* cmask;
    uint32_t* rcnt;    // candidates per row (CAP + 1: overflow, scanned in full)
    uint32_t* ccnt;
    uint32_t* rcand;   // [rows][CAP]
    uint32_t* ccand;
    uint32_t* ovf;     // overflowed items: (side << 31) | launch-local index
    uint32_t* ovf_n;   // [1]
};

// MFMA tile pass, candidate collection, recheck and overflow scans of the
// pairs in `pairs` (all sets safe; tiles of FTILE); writes fwd / bwd.  ev
// (optional, 2 events) are recorded after the tile pass and the collection.
// hi and lo are the two halves of one allocation (the tile kernel addresses
// lo as an offset from hi).
hipError_t launch_filtered(const float* desc, const float* norm, const uint16_t* hi, const uint16_t* lo,
                           const PairDesc* pairs, int npairs, int64_t ntiles, int64_t rows_total, int64_t cols_total,
                           const FiltBufs& fb, Best2* fwd, Best2* bwd, hipStream_t st, hipEvent_t* ev);

// Check of the estimate's error bound: max over all pairs' elements of
// |exact - estimate| / (|a|^2 + |b|^2), as float bits into *fb.ovf_n (which
// the caller zeroes first).
hipError_t launch_filter_check(const float* desc, const float* norm, const uint16_t* hi, const uint16_t* lo,
                               const PairDesc* pairs, int npairs, int64_t ntiles, const FiltBufs& fb,
                               hipStream_t st);

// Exact distance tiles: every (a-tile, b-tile) of every pair; writes the row
// partial (over the tile's 64 b descriptors) of each a descriptor and the
// column partial (over the tile's 64 a descriptors) of each b descriptor.
hipError_t launch_dist_tiles(const float* desc, const PairDesc* pairs, int npairs, int64_t ntiles, Best2* rpart,
                             Best2* cpart, hipStream_t st);

// Merge the partials of every descriptor into fwd / bwd (rows_total /
// cols_total count the launch-local rows / columns).
hipError_t launch_merge_partials(const PairDesc* pairs, int npairs, int64_t rows_total, int64_t cols_total,
                                 const Best2* rpart, const Best2* cpart, Best2* fwd, Best2* bwd, hipStream_t st);

// Mutual check + ratio tests, order-preserving compaction per pair.
hipError_t launch_select(const PairDesc* pairs, int npairs, const Best2* fwd, const Best2* bwd, float ratio,
                         uint32_t* out_a, uint32_t* out_b, int64_t* counts, hipStream_t st);

// Dense copy of the matches: prefix[p] = matches of the pairs before p;
// da / db[prefix[p] ..] = pair p's matches.
hipError_t launch_pack(const PairDesc* pairs, int npairs, const int64_t* counts, int64_t* prefix,
                       const uint32_t* oa, const uint32_t* ob, uint32_t* da, uint32_t* db, hipStream_t st);

}  // namespace mt
