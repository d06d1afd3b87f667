// match.hip -- brute-force descriptor matching of image pairs on the device
// (sfm-lite/src/matching.rs:22-98 match_pair, called per job pair at
// sfm-lite/src/lib.rs:221-224).
//
// Reference semantics kept bit for bit:
//   * distance s(i, j) = sum over k = 0..127 of (a_k - b_k)^2, each difference,
//     square and running sum rounded to f32, in k order (no FMA);
//   * per descriptor the best (smallest s, first index on ties) and the second
//     best (the second order statistic of the multiset of s, +inf if absent);
//     NaN distances are never taken;
//   * a match (i, j) needs j = best(i), i = best(j), finite second(i) and
//     second(j), best(i) <= 0.9, best(i) <= ratio * second(i) and
//     best(j) <= ratio * second(j) (ratio in f32); matches are listed in
//     increasing i.
//
// The reference runs two passes (a against b, b against a) that compute the
// same distance twice (the square hides the sign of a_k - b_k).  Here each
// distance is computed once in a 64 x 64 tile and folded into both the row
// (a-side) and the column (b-side) statistics.  The per-tile statistics are
// merged with an order-free rule: best = lexicographic min of (s, index),
// second = min(max(best1, best2), second1, second2) -- the second order
// statistic of the union -- so tiles and lanes merge in any order and still
// give the sequential scan's answer.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "match_kernels.h"

#pragma clang fp contract(off)

namespace mt {

constexpr int KC = 64;   // descriptor components staged per chunk

__device__ __forceinline__ void push(float v, uint32_t j, float& b, uint32_t& bi, float& s) {
    if (v < b) {
        s = b;
        b = v;
        bi = j;
    } else if (v < s) {
        s = v;
    }
}

__device__ __forceinline__ void merge(float& b, uint32_t& bi, float& s, float b2, uint32_t i2, float s2) {
    const float sec = fminf(fmaxf(b, b2), fminf(s, s2));
    if (b2 < b || (b2 == b && i2 < bi)) {
        b = b2;
        bi = i2;
    }
    s = sec;
}

__device__ __forceinline__ void merge_lanes(float& b, uint32_t& bi, float& s, int m) {
    const float b2 = __shfl_xor(b, m);
    const uint32_t i2 = (uint32_t)__shfl_xor((int)bi, m);
    const float s2 = __shfl_xor(s, m);
    merge(b, bi, s, b2, i2, s2);
}

// Flattened tile t -> pair: the last pair whose tile0 <= t.
// The same search by one wave: tile0 ascends, so the pairs with tile0 <= t
// are a prefix; each round tests 64 of them at once (a handful of
// independent loads instead of log2(npairs) dependent ones).  Every lane of
// the wave must call it.
__device__ __forceinline__ int find_pair_wave(const PairDesc* pairs, int npairs, int64_t t) {
    const int lane = threadIdx.x & 63;
    int best = 0;
    for (int base = 0; base < npairs; base += 64) {
        const int p = base + lane;
        const uint64_t m = __ballot(p < npairs && pairs[p].tile0 <= t);
        if (m) best = base + 63 - __builtin_clzll(m);
        if (m != ~0ull) break;
    }
    return best;
}

__device__ __forceinline__ int find_pair(const PairDesc* pairs, int npairs, int64_t t) {
    int lo = 0, hi = npairs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pairs[mid].tile0 <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Stage KC components of 64 descriptors as [k][d] (d fastest): lane d reads
// 16 bytes of descriptor d, so the LDS writes of a wave hit 64 consecutive
// words.  Rows past the set are zero (their distances are masked later).
__device__ __forceinline__ void stage(float (*s)[TILE], const float* __restrict__ desc, int64_t row0, int n, int kc) {
    const int d = threadIdx.x & 63;
    const bool ok = d < n;
    const float4* src = reinterpret_cast<const float4*>(desc + (row0 + (ok ? d : 0)) * DIM + kc);
#pragma unroll
    for (int it = 0; it < 4; ++it) {
        const int q = (threadIdx.x >> 6) + 4 * it;   // float4 index within the chunk
        float4 v = ok ? src[q] : make_float4(0.f, 0.f, 0.f, 0.f);
        s[q * 4 + 0][d] = v.x;
        s[q * 4 + 1][d] = v.y;
        s[q * 4 + 2][d] = v.z;
        s[q * 4 + 3][d] = v.w;
    }
}

// 256 threads; thread (ty, tx) = (tid / 16, tid % 16) owns rows 4ty..4ty+3 and
// columns 4tx..4tx+3 of the tile.  The logical tile index is spread so that
// consecutive tiles of one pair run on the same XCD (blocks go round-robin
// over the 8 XCDs, each with its own L2).
__global__ __launch_bounds__(256) void dist_tile_kernel(const float* __restrict__ desc,
                                                        const PairDesc* __restrict__ pairs, int npairs,
                                                        int64_t ntiles, int64_t per_xcd, Best2* __restrict__ rpart,
                                                        Best2* __restrict__ cpart) {
    const int64_t t = (int64_t)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (t >= ntiles) return;
    __shared__ int s_pair;
    if (threadIdx.x == 0) s_pair = find_pair(pairs, npairs, t);
    __syncthreads();
    const PairDesc P = pairs[s_pair];
    const int64_t lt = t - P.tile0;
    const int rt = (int)(lt / P.tb), ct = (int)(lt - (int64_t)rt * P.tb);
    const int i0 = rt * TILE, j0 = ct * TILE;
    const int nra = min(TILE, P.na - i0), nrb = min(TILE, P.nb - j0);

    __shared__ float sa[KC][TILE];
    __shared__ float sb[KC][TILE];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    float acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = 0.0f;

    for (int kc = 0; kc < DIM; kc += KC) {
        if (kc) __syncthreads();
        stage(sa, desc, P.a + i0, nra, kc);
        stage(sb, desc, P.b + j0, nrb, kc);
        __syncthreads();
#pragma unroll 4
        for (int k = 0; k < KC; ++k) {
            const float4 av = *reinterpret_cast<const float4*>(&sa[k][ty * 4]);
            const float4 bv = *reinterpret_cast<const float4*>(&sb[k][tx * 4]);
            const float a_[4] = {av.x, av.y, av.z, av.w};
            const float b_[4] = {bv.x, bv.y, bv.z, bv.w};
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const float d = a_[r] - b_[c];
                    const float sq = d * d;
                    acc[r][c] = acc[r][c] + sq;
                }
        }
    }

    // rows (a side): scan own 4 columns in j order, merge over the 16 lanes
    // of the row group, lane tx == 0 writes.
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        float b = INFINITY, s = INFINITY;
        uint32_t bi = NONE;
        const int i = ty * 4 + r;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int j = tx * 4 + c;
            if (j < nrb && i < nra) push(acc[r][c], (uint32_t)(j0 + j), b, bi, s);
        }
        merge_lanes(b, bi, s, 1);
        merge_lanes(b, bi, s, 2);
        merge_lanes(b, bi, s, 4);
        merge_lanes(b, bi, s, 8);
        if (tx == 0 && i < nra) {
            Best2 o;
            o.b = b;
            o.i = bi;
            o.s = s;
            o.pad = 0;
            rpart[P.rpart + (int64_t)ct * P.na + i0 + i] = o;
        }
    }

    // columns (b side): scan own 4 rows in i order, merge over the 4 row
    // groups of the wave, then over the 4 waves through LDS.
    __syncthreads();   // sa / sb are reused below
    float* sred = &sa[0][0];   // [4 waves][64 columns][3]
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float b = INFINITY, s = INFINITY;
        uint32_t bi = NONE;
        const int j = tx * 4 + c;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = ty * 4 + r;
            if (j < nrb && i < nra) push(acc[r][c], (uint32_t)(i0 + i), b, bi, s);
        }
        merge_lanes(b, bi, s, 16);
        merge_lanes(b, bi, s, 32);
        if ((threadIdx.x & 63) < 16) {
            float* o = sred + (w * TILE + j) * 3;
            o[0] = b;
            o[1] = __uint_as_float(bi);
            o[2] = s;
        }
    }
    __syncthreads();
    if (threadIdx.x < TILE && (int)threadIdx.x < nrb) {
        const int j = threadIdx.x;
        float b = sred[j * 3 + 0], s = sred[j * 3 + 2];
        uint32_t bi = __float_as_uint(sred[j * 3 + 1]);
#pragma unroll
        for (int ww = 1; ww < 4; ++ww) {
            const float* o = sred + (ww * TILE + j) * 3;
            merge(b, bi, s, o[0], __float_as_uint(o[1]), o[2]);
        }
        Best2 o;
        o.b = b;
        o.i = bi;
        o.s = s;
        o.pad = 0;
        cpart[P.cpart + (int64_t)rt * P.nb + j0 + j] = o;
    }
}

// Launch-local row item q in [0, rows_total): pair by lrow; columns by lcol.
__device__ __forceinline__ int find_pair_by(const PairDesc* pairs, int npairs, int64_t q, bool cols) {
    int lo = 0, hi = npairs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int64_t base = cols ? pairs[mid].lcol : pairs[mid].lrow;
        if (base <= q)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void merge_partials_kernel(const PairDesc* __restrict__ pairs, int npairs,
                                                             int64_t rows_total, int64_t cols_total,
                                                             const Best2* __restrict__ rpart,
                                                             const Best2* __restrict__ cpart, Best2* __restrict__ fwd,
                                                             Best2* __restrict__ bwd) {
    const bool cols = blockIdx.y == 1;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (cols ? cols_total : rows_total)) return;
    const PairDesc P = pairs[find_pair_by(pairs, npairs, q, cols)];
    const int64_t local = q - (cols ? P.lcol : P.lrow);
    const int nt = cols ? P.ta : P.tb;
    const int n = cols ? P.nb : P.na;
    const Best2* src = (cols ? cpart + P.cpart : rpart + P.rpart) + local;
    float b = INFINITY, s = INFINITY;
    uint32_t bi = NONE;
    for (int t = 0; t < nt; ++t) {
        const Best2 p = src[(int64_t)t * n];
        merge(b, bi, s, p.b, p.i, p.s);
    }
    Best2 o;
    o.b = b;
    o.i = bi;
    o.s = s;
    o.pad = 0;
    (cols ? bwd + P.fcol : fwd + P.frow)[local] = o;
}

// One block per pair: the reference's final loop over a's descriptors,
// compacted in order with a block-wide prefix sum.
__global__ __launch_bounds__(1024) void select_kernel(const PairDesc* __restrict__ pairs,
                                                      const Best2* __restrict__ fwd, const Best2* __restrict__ bwd,
                                                      float ratio, uint32_t* __restrict__ out_a,
                                                      uint32_t* __restrict__ out_b, int64_t* __restrict__ counts) {
    const PairDesc P = pairs[blockIdx.x];
    __shared__ int s_wave[16];
    __shared__ int s_base;
    if (threadIdx.x == 0) s_base = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const bool any = P.na > 0 && P.nb > 0;
    for (int c0 = 0; any && c0 < P.na; c0 += blockDim.x) {
        const int i = c0 + threadIdx.x;
        bool keep = false;
        uint32_t jb = 0;
        if (i < P.na) {
            const Best2 f = fwd[P.frow + i];
            jb = f.i;
            if (f.i != NONE && isfinite(f.s) && !(f.b > 0.9f) && !(f.b > ratio * f.s)) {
                const Best2 g = bwd[P.fcol + f.i];
                keep = g.i == (uint32_t)i && isfinite(g.s) && !(g.b > ratio * g.s);
            }
        }
        const uint64_t m = __ballot(keep);
        const int before = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) s_wave[wv] = __popcll(m);
        __syncthreads();
        int off = s_base;
        for (int q = 0; q < wv; ++q) off += s_wave[q];
        if (keep) {
            out_a[P.out + off + before] = (uint32_t)i;
            out_b[P.out + off + before] = jb;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int tot = 0;
            for (int q = 0; q < nw; ++q) tot += s_wave[q];
            s_base += tot;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) counts[blockIdx.x] = any ? s_base : 0;
}

// --------------------------------------------------------------------------
// Filtered search.
//
// For safe sets (every nonzero component in [2^-40, 2^40]) the distance
// s(i, j) is first estimated from the MFMA product of the bf16 split
// x = hi + lo + r (|lo| <= 2^-8 |x|, |r| <= 2^-17 |x|: bf16 keeps 8
// significant bits):
//   A(i, j) = |a_i|^2 + |b_j|^2 - 2 (hi_a.hi_b + hi_a.lo_b + lo_a.hi_b),
// with |s(i, j) - A(i, j)| <= EPS_REL (|a_i|^2 + |b_j|^2) =: E(i, j).  The
// bound covers the neglected lo.lo and r terms (<= 2^-15 sum |a_k b_k|, so
// <= 2^-15 (|a|^2 + |b|^2) in A), f32 accumulation of the 384 exact bf16
// products and of the norms, the final combination, and the rounding of the
// reference's own sum (<= 130 * 2^-24 |a - b|^2); all together below
// 8e-5 (|a|^2 + |b|^2), a sixth of EPS_REL.  The component range keeps
// every product and partial sum a normal f32, so the relative bounds hold.
//
// Pass 1: per descriptor, T = the second smallest upper bound U = A + E.
//   At least two exact distances are <= T, so the exact second-nearest
//   distance s2 <= T.
// Pass 2: candidates = the descriptors with lower bound L = A - E <= T.
//   Every descriptor whose exact distance is <= s2 -- the nearest (all of
//   them on ties) and the second nearest -- is a candidate.
// Recheck: the exact distances of the candidates, merged with the
//   order-free rule, give the reference's (best index, best, second).
//   Descriptors with more than CAP candidates are scanned in full.
// --------------------------------------------------------------------------

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint16_t bf16_rn(float f) {   // round to nearest even (finite input)
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ bf16x8 as_bf(us8 v) { return __builtin_bit_cast(bf16x8, v); }

// 16 threads per descriptor, 8 components each.
__global__ __launch_bounds__(256) void prep_kernel(const float* __restrict__ desc, int64_t n,
                                                   const int64_t* __restrict__ set_off, int nsets,
                                                   float* __restrict__ norm, uint16_t* __restrict__ hi,
                                                   uint16_t* __restrict__ lo, int* __restrict__ bad) {
    const int64_t d = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
    const int seg = threadIdx.x & 15;
    float ss = 0.0f;
    bool badv = false;
    if (d < n) {
        const float4* src = reinterpret_cast<const float4*>(desc + d * DIM + seg * 8);
        const float4 v0 = src[0], v1 = src[1];
        const float x[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
        us8 h, l;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const float ax = fabsf(x[q]);
            badv |= !(ax == 0.0f || (ax >= 0x1p-40f && ax <= 0x1p40f));
            const uint16_t hb = bf16_rn(x[q]);
            const float hf = __uint_as_float((uint32_t)hb << 16);
            h[q] = hb;
            l[q] = bf16_rn(x[q] - hf);   // x - hf is exact
            ss = ss + x[q] * x[q];
        }
        *reinterpret_cast<us8*>(hi + d * DIM + seg * 8) = h;
        *reinterpret_cast<us8*>(lo + d * DIM + seg * 8) = l;
    }
    ss += __shfl_xor(ss, 1);
    ss += __shfl_xor(ss, 2);
    ss += __shfl_xor(ss, 4);
    ss += __shfl_xor(ss, 8);
    if (d < n && seg == 0) norm[d] = ss;
    if (d < n && badv) {
        int lo_s = 0, hi_s = nsets - 1;   // the last set starting at or before d
        while (lo_s < hi_s) {
            const int mid = (lo_s + hi_s + 1) >> 1;
            if (set_off[mid] <= d)
                lo_s = mid;
            else
                hi_s = mid - 1;
        }
        atomicOr(&bad[lo_s], 1);
    }
}

// (u1 <= u2) <- the two smallest of {u1, u2, v}; med3 is the middle value.
__device__ __forceinline__ void two_min(float& u1, float& u2, float v) {
    u2 = __builtin_amdgcn_fmed3f(u1, u2, v);
    u1 = fminf(u1, v);
}

__device__ __forceinline__ void two_min_merge(float& u1, float& u2, float v1, float v2) {
    u2 = fminf(fmaxf(u1, v1), fminf(u2, v2));
    u1 = fminf(u1, v1);
}

// Lane exchange inside a 16-lane row through DPP (a VALU move, no LDS):
// 0xB1 = quad_perm [1,0,3,2], 0x4E = quad_perm [2,3,0,1], 0x141 =
// row_half_mirror, 0x140 = row_mirror.  The four steps pair every lane with
// each of the other 15 through the merge tree.
template <int CTRL>
__device__ __forceinline__ int dpp_i(int x) {
    return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xF, 0xF, true);
}

// two_min_merge with the partner lane's pair under DPP pattern CTRL, on the
// bit patterns: the upper bounds are >= +0 or +inf (U >= the exact distance
// >= 0), where signed-integer order is float order.  Integer min / max take
// the DPP operand directly (v_min_i32_dpp); float min / max would first
// canonicalize the moved value.
template <int CTRL>
__device__ __forceinline__ void two_min_dpp(int& u1, int& u2) {
    const int a = max(u1, dpp_i<CTRL>(u1));
    const int b = min(u2, dpp_i<CTRL>(u2));
    u1 = min(u1, dpp_i<CTRL>(u1));
    u2 = min(a, b);
}

// One step of a transposed two-min tree over the 16 lanes of a row: the
// lane holds the pairs of 2H rows (bit patterns, as in two_min_dpp), keeps
// rows [0, H) or [H, 2H) (hi), sends the other half to its partner under
// DPP pattern CTRL (which has the opposite hi and the same kept rows) and
// merges the partner's half into its own.  The kept pairs end in a, b [0, H).
template <int CTRL, int H>
__device__ __forceinline__ void two_min_tstep(int* a, int* b, bool hi) {
#pragma unroll
    for (int t = 0; t < H; ++t) {
        const int ka = hi ? a[t + H] : a[t], kb = hi ? b[t + H] : b[t];
        const int sa = hi ? a[t] : a[t + H], sb = hi ? b[t] : b[t + H];
        const int pa = dpp_i<CTRL>(sa), pb = dpp_i<CTRL>(sb);
        a[t] = min(ka, pa);
        b[t] = min(max(ka, pa), min(kb, pb));
    }
}

// The 16 pairs a lane holds (row j = 0..15 of its 16-lane row group), each
// merged over the 16 lanes of the group: row_mirror pairs lane r with
// 15 - r, row_half_mirror then r with r ^ 7, the quad permutes with r ^ 2
// and r ^ 1, so lane r ends with row j = r, merged over all 16 lanes, in
// 15 merges instead of 16 x 4.
__device__ __forceinline__ void two_min_rows16(int* a, int* b, int r16) {
    two_min_tstep<0x140, 8>(a, b, (r16 & 8) != 0);
    two_min_tstep<0x141, 4>(a, b, (r16 & 4) != 0);
    two_min_tstep<0x4E, 2>(a, b, (r16 & 2) != 0);
    two_min_tstep<0xB1, 1>(a, b, (r16 & 1) != 0);
}

// two_min_merge of lane l's pair with lane (l ^ 16)'s (SPAN 16) or
// (l ^ 32)'s (SPAN 32): a permlane swap of a register with a copy of itself
// leaves each lane holding its own value and its partner's, in some order,
// and the merge is symmetric.
template <int SPAN>
__device__ __forceinline__ void two_min_xlane(float& u1, float& u2) {
    const int a = __float_as_int(u1), b = __float_as_int(u2);
    const auto sa = SPAN == 16 ? __builtin_amdgcn_permlane16_swap(a, a, false, false)
                               : __builtin_amdgcn_permlane32_swap(a, a, false, false);
    const auto sb = SPAN == 16 ? __builtin_amdgcn_permlane16_swap(b, b, false, false)
                               : __builtin_amdgcn_permlane32_swap(b, b, false, false);
    u1 = __int_as_float(sa[0]);
    u2 = __int_as_float(sb[0]);
    two_min_merge(u1, u2, __int_as_float(sa[1]), __int_as_float(sb[1]));
}

// LDS layout of a staged chunk: [128 rows][4 segments of 8 components],
// 64 bytes a row with no padding; segment s of row R sits in slot
// s ^ ((R >> 2) & 3), so the 16 lanes that read segment kq of 16 consecutive
// rows hit 16 distinct 16-byte bank groups (all 64 banks once).
__device__ __forceinline__ int lds_slot(int row, int seg) { return row * 32 + ((seg ^ ((row >> 2) & 3)) << 3); }

// PASS 1: per row and per column of a 128 x 128 tile, the two smallest
//         upper bounds U (u1 <= u2), a lower bound of the tile's L, and a
//         128-bit mask of the elements whose L is <= u2.  The row's global
//         threshold T (its second smallest U over all tiles) is <= u2 of
//         every tile, so every element with L <= T is in a mask.
// PASS 3: (check) max over the tile of |exact - estimate| / (|a|^2 + |b|^2)
//         into *fb.ovf_n as float bits.
// 4 waves in a 2 x 2 grid, each 64 x 64 = 4 x 4 MFMA 16x16x32 fragments.
// Three workgroups per CU (34 KB of LDS each, which would allow four) need
// <= 168 registers per lane: the compiler then keeps two lane coordinates in scratch across the
// product loop (nothing inside it), which costs far less than the third
// wave per SIMD gains (1.11 -> 0.89 ms per batch of 190 pairs of 2048
// descriptors, with the norms written to LDS before the loop).  The budget
// has no slack: changes to the loop's address arithmetic have made the
// allocator spill an accumulator inside the loop.
template <int PASS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8))) void filt_tile_kernel(const float* __restrict__ desc,
                                                        const float* __restrict__ norm,
                                                        const uint16_t* __restrict__ hi,
                                                        const uint16_t* __restrict__ lo,
                                                        const PairDesc* __restrict__ pairs, int npairs,
                                                        int64_t ntiles, int64_t per_xcd, FiltBufs fb) {
    const int64_t t = (int64_t)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (t >= ntiles) return;
    __shared__ int s_pair;
    if (threadIdx.x < 64) {
        const int p = find_pair_wave(pairs, npairs, t);
        if (threadIdx.x == 0) s_pair = p;
    }
    if (PASS == 1 && t == 0 && threadIdx.x == 0) *fb.ovf_n = 0u;   // read by the collection after this launch
    __syncthreads();
    const PairDesc P = pairs[s_pair];
    const int64_t lt = t - P.tile0;
    const int rt = (int)(lt / P.tb), ct = (int)(lt - (int64_t)rt * P.tb);
    const int i0 = rt * FTILE, j0 = ct * FTILE;
    const int nra = min(FTILE, P.na - i0), nrb = min(FTILE, P.nb - j0);
    // squared norms of the tile's rows (threads 0..127) and columns (128..255),
    // pre-scaled by 1 + EPS_REL in pass 1; rows / columns past a set hold +inf
    // in pass 1, 0 in pass 3.  Written to LDS before the products (the
    // product loop's first barrier orders them), so that no register holds
    // them across the loop.
    __shared__ float s_norm[2 * FTILE];
    {
        const int e = threadIdx.x & (FTILE - 1);
        const bool is_row = threadIdx.x < FTILE;
        float my_norm = PASS == 1 ? INFINITY : 0.0f;
        if (e < (is_row ? nra : nrb))
            my_norm = (PASS == 1 ? 1.0f + EPS_REL : 1.0f) * norm[is_row ? P.a + i0 + e : P.b + j0 + e];
        s_norm[threadIdx.x] = my_norm;
    }

    __shared__ __align__(16) uint16_t sAh[FTILE * 32], sAl[FTILE * 32], sBh[FTILE * 32], sBl[FTILE * 32];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, wr = w >> 1, wc = w & 1;
    const int r16 = lane & 15, kq = lane >> 4;
    f32x4 acc[4][4];
#pragma unroll
    for (int mf = 0; mf < 4; ++mf)
#pragma unroll
        for (int nf = 0; nf < 4; ++nf) acc[mf][nf] = f32x4{0.f, 0.f, 0.f, 0.f};

    // Staging: each chunk of 32 components goes from memory straight into
    // LDS (global_load_lds_dwordx4: lane L of a wave writes the 16 bytes at
    // the wave's LDS base + 16 L).  A wave's instruction fills 16 rows of one
    // array, so the lane loads the segment that belongs in slot L & 3 of row
    // L >> 2.  Rows past a set re-read its last row (in bounds; masked in the
    // epilogue).  A chunk's fragments are read into registers first; after a
    // barrier the next chunk's loads refill the buffer while the MFMAs run,
    // with no staging registers.
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const uint16_t* ga[2];
    const uint16_t* gb[2];
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const int row = (it * 4 + wv) * 16 + (lane >> 2);
        const int gseg = (lane & 3) ^ ((row >> 2) & 3);
        ga[it] = hi + (size_t)(P.a + i0 + min(row, nra - 1)) * DIM + gseg * 8;
        gb[it] = hi + (size_t)(P.b + j0 + min(row, nrb - 1)) * DIM + gseg * 8;
    }
    const ptrdiff_t lo_off = lo - hi;
    typedef __attribute__((address_space(3))) void* lds_ptr;
    auto issue = [&](int kc) {
#pragma unroll
        for (int it = 0; it < 2; ++it) {
            const int base = (it * 4 + wv) * 16 * 32;   // first u16 of the wave's 16 rows
            __builtin_amdgcn_global_load_lds((const void*)(ga[it] + kc), (lds_ptr)(sAh + base), 16, 0, 0);
            __builtin_amdgcn_global_load_lds((const void*)(ga[it] + lo_off + kc), (lds_ptr)(sAl + base), 16, 0, 0);
            __builtin_amdgcn_global_load_lds((const void*)(gb[it] + kc), (lds_ptr)(sBh + base), 16, 0, 0);
            __builtin_amdgcn_global_load_lds((const void*)(gb[it] + lo_off + kc), (lds_ptr)(sBl + base), 16, 0, 0);
        }
    };
    issue(0);
#pragma unroll
    for (int kc = 0; kc < DIM; kc += 32) {
        // the previous chunk's MFMAs stay between its loads' issue and this wait
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_waitcnt(0);   // this wave's loads of the chunk have landed
        __syncthreads();                 // and every other wave's
        us8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            const int row = wr * 64 + f * 16 + r16, col = wc * 64 + f * 16 + r16;
            ah[f] = *reinterpret_cast<const us8*>(sAh + lds_slot(row, kq));
            al[f] = *reinterpret_cast<const us8*>(sAl + lds_slot(row, kq));
            bh[f] = *reinterpret_cast<const us8*>(sBh + lds_slot(col, kq));
            bl[f] = *reinterpret_cast<const us8*>(sBl + lds_slot(col, kq));
        }
        __syncthreads();   // every wave holds its fragments: the buffer may be refilled
        if (kc + 32 < DIM) issue(kc + 32);
        // the loads go out before the MFMAs (the scheduler would otherwise
        // sink them to the end of the chunk, where nothing hides them)
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int mf = 0; mf < 4; ++mf)
#pragma unroll
            for (int nf = 0; nf < 4; ++nf) {
                acc[mf][nf] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(ah[mf]), as_bf(bh[nf]), acc[mf][nf], 0, 0, 0);
                acc[mf][nf] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(ah[mf]), as_bf(bl[nf]), acc[mf][nf], 0, 0, 0);
                acc[mf][nf] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(al[mf]), as_bf(bh[nf]), acc[mf][nf], 0, 0, 0);
            }
    }

    // element (mf, nf, i) of this lane: tile row R = wr*64 + mf*16 + kq*4 + i,
    // tile column C = wc*64 + nf*16 + r16
    // Norms, pre-scaled by 1 + EPS_REL in pass 1: tU = (1 + EPS)(|a|^2 +
    // |b|^2), U = tU - 2G, L = U - k tU with k = 2 EPS / (1 + EPS).  Rows /
    // columns past a set get +inf: U = +inf never enters the two smallest and
    // L = NaN fails every compare.
    float nrow[4][4], ncol[4];
#pragma unroll
    for (int mf = 0; mf < 4; ++mf)
#pragma unroll
        for (int i = 0; i < 4; ++i) nrow[mf][i] = s_norm[wr * 64 + mf * 16 + kq * 4 + i];
#pragma unroll
    for (int nf = 0; nf < 4; ++nf) ncol[nf] = s_norm[FTILE + wc * 64 + nf * 16 + r16];

    if constexpr (PASS == 1) {
        float ru1[4][4], ru2[4][4], cu1[4], cu2[4];
#pragma unroll
        for (int mf = 0; mf < 4; ++mf)
#pragma unroll
            for (int i = 0; i < 4; ++i) ru1[mf][i] = ru2[mf][i] = INFINITY;
#pragma unroll
        for (int nf = 0; nf < 4; ++nf) cu1[nf] = cu2[nf] = INFINITY;
#pragma unroll
        for (int mf = 0; mf < 4; ++mf)
#pragma unroll
            for (int nf = 0; nf < 4; ++nf)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float U = __builtin_fmaf(-2.0f, acc[mf][nf][i], nrow[mf][i] + ncol[nf]);
                    two_min(ru1[mf][i], ru2[mf][i], U);
                    two_min(cu1[nf], cu2[nf], U);
                }
        // rows: over the 16 lanes of a kq group (lane r16 ends with row
        // mf = r16 / 4, i = r16 % 4), then the two column waves
        int ra[16], rb[16];
#pragma unroll
        for (int mf = 0; mf < 4; ++mf)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                ra[mf * 4 + i] = __float_as_int(ru1[mf][i]);
                rb[mf * 4 + i] = __float_as_int(ru2[mf][i]);
            }
        two_min_rows16(ra, rb, r16);
        // columns: over the 4 kq groups, then the two row waves
#pragma unroll
        for (int nf = 0; nf < 4; ++nf) {
            two_min_xlane<16>(cu1[nf], cu2[nf]);
            two_min_xlane<32>(cu1[nf], cu2[nf]);
        }
        __syncthreads();   // the staging buffers are reused below
        float2* srow = reinterpret_cast<float2*>(sAh);               // [2 column waves][128 rows]
        float2* scol = reinterpret_cast<float2*>(sAh) + 2 * FTILE;   // [2 row waves][128 columns]
        float2* sur = reinterpret_cast<float2*>(sAl);                // merged (u1, u2) per row
        float2* suc = reinterpret_cast<float2*>(sAl) + FTILE;        // per column
        uint32_t* srm = reinterpret_cast<uint32_t*>(sBh);            // [128 rows][4 words]: columns with L <= u2
        uint32_t* scm = srm + 4 * FTILE;                             // [128 columns][4 words]: rows
        uint32_t* snmax = reinterpret_cast<uint32_t*>(sBl);          // [2]: max scaled norm of rows, columns
        srow[wc * FTILE + wr * 64 + (r16 >> 2) * 16 + kq * 4 + (r16 & 3)] =
            make_float2(__int_as_float(ra[0]), __int_as_float(rb[0]));
        if (kq == 0) {
#pragma unroll
            for (int nf = 0; nf < 4; ++nf) {
                const int C = wc * 64 + nf * 16 + r16;
                scol[wr * FTILE + C] = make_float2(cu1[nf], cu2[nf]);
            }
        }
        if (threadIdx.x < 2) snmax[threadIdx.x] = 0u;
        __syncthreads();
        {
            const int q = threadIdx.x, e = q & (FTILE - 1);
            const bool is_row = q < FTILE;
            float2 a = is_row ? srow[e] : scol[e];
            const float2 b = is_row ? srow[FTILE + e] : scol[FTILE + e];
            two_min_merge(a.x, a.y, b.x, b.y);
            (is_row ? sur : suc)[e] = a;
            // scaled norms are >= 0, so their bit patterns order as the values;
            // a wave holds 64 rows or 64 columns: its maximum, then one atomic
            // per wave (256 atomics on two words serialised in the LDS)
            unsigned nm = e < (is_row ? nra : nrb) ? __float_as_uint(s_norm[q]) : 0u;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) nm = max(nm, (unsigned)__shfl_xor((int)nm, off));
            if ((threadIdx.x & 63) == 0) atomicMax(&snmax[is_row ? 0 : 1], nm);
        }
        __syncthreads();
        // Second sweep over the accumulators: bit masks of the elements with
        // L <= the tile's u2, per row (columns) and per column (rows).  The
        // empty asm makes the accumulators opaque, so tU and U are recomputed
        // from the accumulator registers instead of being kept live (64 more
        // registers) across the reductions; the norms are read again from
        // LDS (the same values) rather than held across them, and the row /
        // column bases are made opaque too, so that the row indices are
        // recomputed here rather than kept from the first sweep.
#pragma unroll
        for (int mf = 0; mf < 4; ++mf)
#pragma unroll
            for (int nf = 0; nf < 4; ++nf) asm volatile("" : "+a"(acc[mf][nf]));
        int rbase = wr * 64 + kq * 4, cbase = wc * 64 + r16;
        asm volatile("" : "+v"(rbase), "+v"(cbase));
#pragma unroll
        for (int mf = 0; mf < 4; ++mf)
#pragma unroll
            for (int i = 0; i < 4; ++i) nrow[mf][i] = s_norm[rbase + mf * 16 + i];
#pragma unroll
        for (int nf = 0; nf < 4; ++nf) ncol[nf] = s_norm[FTILE + cbase + nf * 16];
        const float kb = 2.0f * EPS_REL / (1.0f + EPS_REL);
        float tc[4];
        uint32_t cw[4][2];   // per column block: rows wr*64 + (0..31), wr*64 + (32..63)
#pragma unroll
        for (int nf = 0; nf < 4; ++nf) {
            tc[nf] = suc[cbase + nf * 16].y;
            cw[nf][0] = cw[nf][1] = 0u;
        }
#pragma unroll
        for (int mf = 0; mf < 4; ++mf)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int R = rbase + mf * 16 + i;
                const float tr = sur[R].y;
                // the row test's wave mask holds, for each of the four rows
                // of this (mf, i) (kq = 0..3), its 16 columns of block nf at
                // bits kq*16 .. kq*16+15: no lane-by-lane OR is needed
                uint64_t bl[4];
#pragma unroll
                for (int nf = 0; nf < 4; ++nf) {
                    const float tU = nrow[mf][i] + ncol[nf];
                    const float U = __builtin_fmaf(-2.0f, acc[mf][nf][i], tU);
                    const float L = __builtin_fmaf(-kb, tU, U);
                    bl[nf] = __ballot(L <= tr);
                    cw[nf][mf >> 1] |= L <= tc[nf] ? 1u << ((mf & 1) * 16 + kq * 4 + i) : 0u;
                }
                // this lane's row: the 16-bit fields of blocks 0, 1 (columns
                // wc*64 + 0..31) and 2, 3 (32..63), paired by a byte permute
                const uint32_t rw0 = __builtin_amdgcn_perm((uint32_t)(bl[1] >> (16 * kq)),
                                                           (uint32_t)(bl[0] >> (16 * kq)), 0x05040100u);
                const uint32_t rw1 = __builtin_amdgcn_perm((uint32_t)(bl[3] >> (16 * kq)),
                                                           (uint32_t)(bl[2] >> (16 * kq)), 0x05040100u);
                // lane r16 == 0 owns the row's two words of this column half
                // (no other wave writes them)
                if (r16 == 0) {
                    srm[R * 4 + wc * 2 + 0] = rw0;
                    srm[R * 4 + wc * 2 + 1] = rw1;
                }
            }
        // columns: OR over the 4 kq groups; lane kq == 0 owns the column's two
        // words of this row half
#pragma unroll
        for (int nf = 0; nf < 4; ++nf) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int v = (int)cw[nf][h];
                const auto s16 = __builtin_amdgcn_permlane16_swap(v, v, false, false);
                const int v2 = s16[0] | s16[1];
                const auto s32 = __builtin_amdgcn_permlane32_swap(v2, v2, false, false);
                cw[nf][h] = (uint32_t)(s32[0] | s32[1]);
            }
            if (kq == 0) {
                const int C = wc * 64 + nf * 16 + r16;
                scm[C * 4 + wr * 2 + 0] = cw[nf][0];
                scm[C * 4 + wr * 2 + 1] = cw[nf][1];
            }
        }
        __syncthreads();
        // records: threads 0..127 rows, 128..255 columns (the thread index
        // made opaque, so that e is recomputed here rather than kept, or
        // spilled, from the top of the kernel)
        int q = threadIdx.x;
        asm volatile("" : "+v"(q));
        const int e = q & (FTILE - 1);
        const bool is_row = q < FTILE;
        if (e < (is_row ? nra : nrb)) {
            const float2 u = is_row ? sur[e] : suc[e];
            const float own = s_norm[q];
            const float other = __uint_as_float(snmax[is_row ? 1 : 0]);
            // every element's L >= u1 - kb * max tU; the factor and the
            // constant cover the rounding of L
            TileMeta m;
            m.u1 = u.x;
            m.u2 = u.y;
            m.lb = u.x - kb * (own + other) * (1.0f + 0x1p-20f) - 0x1p-100f;
            m.pad = 0u;
            const uint32_t* w = (is_row ? srm : scm) + e * 4;
            const uint4 mask = make_uint4(w[0], w[1], w[2], w[3]);
            if (is_row) {
                fb.rmeta[P.rpart + (int64_t)ct * P.na + i0 + e] = m;
                fb.rmask[P.rpart + (int64_t)ct * P.na + i0 + e] = mask;
            } else {
                fb.cmeta[P.cpart + (int64_t)rt * P.nb + j0 + e] = m;
                fb.cmask[P.cpart + (int64_t)rt * P.nb + j0 + e] = mask;
            }
        }
    } else {
        float worst = 0.0f;
#pragma unroll
        for (int mf = 0; mf < 4; ++mf)
#pragma unroll
            for (int nf = 0; nf < 4; ++nf)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int R = wr * 64 + mf * 16 + kq * 4 + i, C = wc * 64 + nf * 16 + r16;
                    if (R >= nra || C >= nrb) continue;
                    const float* x = desc + (P.a + i0 + R) * DIM;
                    const float* y = desc + (P.b + j0 + C) * DIM;
                    float s = 0.0f;
                    for (int k = 0; k < DIM; ++k) {
                        const float d = x[k] - y[k];
                        s = s + d * d;
                    }
                    const float t2 = nrow[mf][i] + ncol[nf];
                    const float A = t2 - 2.0f * acc[mf][nf][i];
                    const float e = t2 > 0.0f ? fabsf(s - A) / t2 : fabsf(s - A);
                    worst = fmaxf(worst, e);
                }
        atomicMax(fb.ovf_n, __float_as_uint(worst));   // non-negative floats order as their bits
    }
}

// Launch-local row / column q: T = the second smallest upper bound over its
// tiles; every element with L <= T lies in a tile whose lower bound is <= T
// and is set in that tile's mask (L <= T <= the tile's u2).  The masked
// elements of those tiles are the candidates -- a superset, which the exact
// recheck cannot be misled by: an extra element has exact distance >= L > T,
// above the second nearest.  More than CAP: the row is scanned in full.
__global__ __launch_bounds__(256) void filt_collect_kernel(const PairDesc* __restrict__ pairs, int npairs,
                                                           int64_t rows_total, int64_t cols_total, FiltBufs fb) {
    const bool cols = blockIdx.y == 1;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (cols ? cols_total : rows_total)) return;
    const PairDesc P = pairs[find_pair_by(pairs, npairs, q, cols)];
    const int64_t local = q - (cols ? P.lcol : P.lrow);
    const int nt = cols ? P.ta : P.tb;
    const int n = cols ? P.nb : P.na;
    const int64_t base = (cols ? P.cpart : P.rpart) + local;
    const TileMeta* meta = (cols ? fb.cmeta : fb.rmeta) + base;
    const uint4* mask = (cols ? fb.cmask : fb.rmask) + base;
    // the lower bounds of the first 16 tiles (2048 columns) stay in
    // registers from the first read of their records
    constexpr int NCACHE = 16;
    float lbs[NCACHE];
    float u1 = INFINITY, u2 = INFINITY;
#pragma unroll
    for (int t = 0; t < NCACHE; ++t) {
        lbs[t] = NAN;
        if (t < nt) {
            const TileMeta m = meta[(int64_t)t * n];
            two_min_merge(u1, u2, m.u1, m.u2);
            lbs[t] = m.lb;
        }
    }
    for (int t = NCACHE; t < nt; ++t) {
        const TileMeta m = meta[(int64_t)t * n];
        two_min_merge(u1, u2, m.u1, m.u2);
    }
    uint32_t c = 0;
    uint32_t* cand = (cols ? fb.ccand : fb.rcand) + (size_t)q * CAP;
    auto take = [&](int t) {
        const uint4 m4 = mask[(int64_t)t * n];
        const uint32_t w[4] = {m4.x, m4.y, m4.z, m4.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            uint32_t bits = w[k];
            while (bits) {
                const int b = __builtin_ctz(bits);
                bits &= bits - 1u;
                if (c < (uint32_t)CAP) cand[c] = (uint32_t)(t * FTILE + k * 32 + b);
                ++c;
            }
        }
    };
    // tiles in increasing order, as one loop over all of them would take them
    uint32_t pass = 0;
#pragma unroll
    for (int t = 0; t < NCACHE; ++t) pass |= lbs[t] <= u2 ? 1u << t : 0u;   // NaN (no tile) fails
    while (pass) {
        const int t = __builtin_ctz(pass);
        pass &= pass - 1u;
        take(t);
    }
    for (int t = NCACHE; t < nt; ++t)
        if (meta[(int64_t)t * n].lb <= u2) take(t);
    if (c > (uint32_t)CAP) {
        c = CAP + 1;
        fb.ovf[atomicAdd(fb.ovf_n, 1u)] = (cols ? 0x80000000u : 0u) | (uint32_t)q;
    }
    (cols ? fb.ccnt : fb.rcnt)[q] = c;
}

// Exact distance of a-side descriptor x and b-side descriptor y, as the
// reference sums it.
__device__ __forceinline__ float exact_dist(const float* __restrict__ x, const float* __restrict__ y) {
    const float4* x4 = reinterpret_cast<const float4*>(x);
    const float4* y4 = reinterpret_cast<const float4*>(y);
    float s = 0.0f;
#pragma unroll 4
    for (int k = 0; k < DIM / 4; ++k) {
        const float4 a = x4[k], b = y4[k];
        float d = a.x - b.x;
        s = s + d * d;
        d = a.y - b.y;
        s = s + d * d;
        d = a.z - b.z;
        s = s + d * d;
        d = a.w - b.w;
        s = s + d * d;
    }
    return s;
}

// The same with one descriptor in LDS: (own - other)^2 has the bits of
// (other - own)^2, so which side is a does not matter.
__device__ __forceinline__ float exact_dist_own(const float4* __restrict__ own, const float* __restrict__ other) {
    const float4* y4 = reinterpret_cast<const float4*>(other);
    float s = 0.0f;
#pragma unroll 4
    for (int k = 0; k < DIM / 4; ++k) {
        const float4 a = own[k], b = y4[k];
        float d = a.x - b.x;
        s = s + d * d;
        d = a.y - b.y;
        s = s + d * d;
        d = a.z - b.z;
        s = s + d * d;
        d = a.w - b.w;
        s = s + d * d;
    }
    return s;
}

// 8 lanes per launch-local row or column: lane s rechecks candidates s and
// s + 8; candidates may come in any order (the merge is order-free).
__global__ __launch_bounds__(256) void filt_recheck_kernel(const float* __restrict__ desc,
                                                           const PairDesc* __restrict__ pairs, int npairs,
                                                           int64_t rows_total, int64_t cols_total, FiltBufs fb,
                                                           Best2* __restrict__ fwd, Best2* __restrict__ bwd) {
    static_assert(CAP == 16, "8 lanes, two candidate slots each");
    // the item's own descriptor, staged once for its 8 lanes (rows padded by
    // one float4 so the 8 items of a wave read different banks)
    __shared__ float4 s_own[32][DIM / 4 + 1];
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t item = g >> 3;
    const int s = (int)(g & 7), li = (int)(threadIdx.x >> 3);
    const bool valid = item < rows_total + cols_total;
    const bool row = item < rows_total;
    const int64_t q = row ? item : item - rows_total;
    float b = INFINITY, sec = INFINITY;
    uint32_t bi = NONE;
    uint32_t cnt = 0;
    PairDesc P;
    int64_t local = 0;
    if (valid) {
        P = pairs[find_pair_by(pairs, npairs, q, !row)];
        local = q - (row ? P.lrow : P.lcol);
        cnt = row ? fb.rcnt[q] : fb.ccnt[q];
        const float4* own = reinterpret_cast<const float4*>(desc + (row ? P.a + local : P.b + local) * DIM);
#pragma unroll
        for (int k = 0; k < 4; ++k) s_own[li][s + 8 * k] = own[s + 8 * k];
    }
    __syncthreads();
    if (valid) {
        for (uint32_t k = (uint32_t)s; cnt <= (uint32_t)CAP && k < cnt; k += 8) {
            const uint32_t other = row ? fb.rcand[(size_t)q * CAP + k] : fb.ccand[(size_t)q * CAP + k];
            const float* y = desc + (row ? P.b + other : P.a + other) * DIM;
            merge(b, bi, sec, exact_dist_own(s_own[li], y), other, INFINITY);
        }
    }
    merge_lanes(b, bi, sec, 1);
    merge_lanes(b, bi, sec, 2);
    merge_lanes(b, bi, sec, 4);
    if (valid && s == 0 && cnt <= (uint32_t)CAP) {
        Best2 o;
        o.b = b;
        o.i = bi;
        o.s = sec;
        o.pad = 0;
        (row ? fwd + P.frow : bwd + P.fcol)[local] = o;
    }
}

// One wave per overflowed row / column: full exact scan.
__global__ __launch_bounds__(256) void filt_overflow_kernel(const float* __restrict__ desc,
                                                            const PairDesc* __restrict__ pairs, int npairs,
                                                            FiltBufs fb, Best2* __restrict__ fwd,
                                                            Best2* __restrict__ bwd) {
    const uint32_t n = *fb.ovf_n;
    const int lane = threadIdx.x & 63;
    const uint32_t nw = gridDim.x * (blockDim.x >> 6);
    for (uint32_t o = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); o < n; o += nw) {
        const uint32_t it = fb.ovf[o];
        const bool row = (it >> 31) == 0;
        const int64_t q = it & 0x7fffffffu;
        const PairDesc P = pairs[find_pair_by(pairs, npairs, q, !row)];
        const int64_t local = q - (row ? P.lrow : P.lcol);
        const int n_other = row ? P.nb : P.na;
        float b = INFINITY, sec = INFINITY;
        uint32_t bi = NONE;
        for (int j = lane; j < n_other; j += 64) {
            const float* x = desc + (row ? P.a + local : P.a + j) * DIM;
            const float* y = desc + (row ? P.b + j : P.b + local) * DIM;
            push(exact_dist(x, y), (uint32_t)j, b, bi, sec);
        }
        for (int m = 1; m < 64; m <<= 1) merge_lanes(b, bi, sec, m);
        if (lane == 0) {
            Best2 ob;
            ob.b = b;
            ob.i = bi;
            ob.s = sec;
            ob.pad = 0;
            (row ? fwd + P.frow : bwd + P.fcol)[local] = ob;
        }
    }
}

// Exclusive prefix of the per-pair match counts (one block), then every
// pair's matches gathered from its slots into one dense list per side.
__global__ __launch_bounds__(1024) void count_scan_kernel(const int64_t* __restrict__ counts, int npairs,
                                                          int64_t* __restrict__ prefix) {
    __shared__ int64_t s_part[1024];
    __shared__ int64_t s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (int c0 = 0; c0 < npairs; c0 += 1024) {
        const int p = c0 + threadIdx.x;
        const int64_t v = p < npairs ? counts[p] : 0;
        s_part[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {   // Hillis-Steele inclusive scan
            const int64_t add = threadIdx.x >= (unsigned)off ? s_part[threadIdx.x - off] : 0;
            __syncthreads();
            s_part[threadIdx.x] += add;
            __syncthreads();
        }
        if (p < npairs) prefix[p] = s_carry + s_part[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) s_carry += s_part[1023];
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void pack_kernel(const PairDesc* __restrict__ pairs,
                                                   const int64_t* __restrict__ counts,
                                                   const int64_t* __restrict__ prefix,
                                                   const uint32_t* __restrict__ oa, const uint32_t* __restrict__ ob,
                                                   uint32_t* __restrict__ da, uint32_t* __restrict__ db) {
    const int p = blockIdx.x;
    const int64_t n = counts[p], src = pairs[p].out, dst = prefix[p];
    for (int64_t k = threadIdx.x; k < n; k += blockDim.x) {
        da[dst + k] = oa[src + k];
        db[dst + k] = ob[src + k];
    }
}

hipError_t launch_pack(const PairDesc* pairs, int npairs, const int64_t* counts, int64_t* prefix,
                       const uint32_t* oa, const uint32_t* ob, uint32_t* da, uint32_t* db, hipStream_t st) {
    if (npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(count_scan_kernel, dim3(1), dim3(1024), 0, st, counts, npairs, prefix);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(pack_kernel, dim3(npairs), dim3(256), 0, st, pairs, counts, prefix, oa, ob, da, db);
    return hipGetLastError();
}

hipError_t launch_dist_tiles(const float* desc, const PairDesc* pairs, int npairs, int64_t ntiles, Best2* rpart,
                             Best2* cpart, hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    const int64_t per_xcd = (ntiles + 7) / 8;
    hipLaunchKernelGGL(dist_tile_kernel, dim3((unsigned)(per_xcd * 8)), dim3(256), 0, st, desc, pairs, npairs, ntiles,
                       per_xcd, rpart, cpart);
    return hipGetLastError();
}

hipError_t launch_merge_partials(const PairDesc* pairs, int npairs, int64_t rows_total, int64_t cols_total,
                                 const Best2* rpart, const Best2* cpart, Best2* fwd, Best2* bwd, hipStream_t st) {
    const int64_t n = rows_total > cols_total ? rows_total : cols_total;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(merge_partials_kernel, dim3((unsigned)((n + 255) / 256), 2), dim3(256), 0, st, pairs, npairs,
                       rows_total, cols_total, rpart, cpart, fwd, bwd);
    return hipGetLastError();
}

hipError_t launch_select(const PairDesc* pairs, int npairs, const Best2* fwd, const Best2* bwd, float ratio,
                         uint32_t* out_a, uint32_t* out_b, int64_t* counts, hipStream_t st) {
    if (npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(select_kernel, dim3(npairs), dim3(1024), 0, st, pairs, fwd, bwd, ratio, out_a, out_b, counts);
    return hipGetLastError();
}

hipError_t launch_prep(const float* desc, int64_t n, const int64_t* set_off, int nsets, float* norm, uint16_t* hi,
                       uint16_t* lo, int* bad, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(prep_kernel, dim3((unsigned)((n + 15) / 16)), dim3(256), 0, st, desc, n, set_off, nsets, norm,
                       hi, lo, bad);
    return hipGetLastError();
}

hipError_t launch_filtered(const float* desc, const float* norm, const uint16_t* hi, const uint16_t* lo,
                           const PairDesc* pairs, int npairs, int64_t ntiles, int64_t rows_total, int64_t cols_total,
                           const FiltBufs& fb, Best2* fwd, Best2* bwd, hipStream_t st, hipEvent_t* ev) {
    if (ntiles <= 0) return hipSuccess;
    const int64_t per_xcd = (ntiles + 7) / 8;
    hipError_t e;
    hipLaunchKernelGGL(filt_tile_kernel<1>, dim3((unsigned)(per_xcd * 8)), dim3(256), 0, st, desc, norm, hi, lo,
                       pairs, npairs, ntiles, per_xcd, fb);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (ev) (void)hipEventRecord(ev[0], st);
    const int64_t n = rows_total > cols_total ? rows_total : cols_total;
    hipLaunchKernelGGL(filt_collect_kernel, dim3((unsigned)((n + 255) / 256), 2), dim3(256), 0, st, pairs, npairs,
                       rows_total, cols_total, fb);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (ev) (void)hipEventRecord(ev[1], st);
    const int64_t items = (rows_total + cols_total) * 8;
    hipLaunchKernelGGL(filt_recheck_kernel, dim3((unsigned)((items + 255) / 256)), dim3(256), 0, st, desc, pairs,
                       npairs, rows_total, cols_total, fb, fwd, bwd);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(filt_overflow_kernel, dim3(256), dim3(256), 0, st, desc, pairs, npairs, fb, fwd, bwd);
    return hipGetLastError();
}

hipError_t launch_filter_check(const float* desc, const float* norm, const uint16_t* hi, const uint16_t* lo,
                               const PairDesc* pairs, int npairs, int64_t ntiles, const FiltBufs& fb,
                               hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    const int64_t per_xcd = (ntiles + 7) / 8;
    hipLaunchKernelGGL(filt_tile_kernel<3>, dim3((unsigned)(per_xcd * 8)), dim3(256), 0, st, desc, norm, hi, lo,
                       pairs, npairs, ntiles, per_xcd, fb);
    return hipGetLastError();
}

}  // namespace mt
