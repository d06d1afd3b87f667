// fuse.hip -- geometric-consistency fusion and voxel dedup on gfx950.
//
// Reference: mvs/src/fuse.rs:32-215 (sample_depth, fuse_depth_maps).
//
// Stages (all on the context's stream):
//   1. candidates: the pixels with a depth of every view that has a depth
//      map (global index = view-major raster order, the reference's merge
//      order) are listed in order by a flag pass and a compaction; one
//      thread per listed pixel: f64 back-projection, projection into every
//      view of its fusion order, f32 bilinear depth sample, relative
//      agreement test, average.  Writes a flag, the f64 point and the
//      colour per pixel.  (Only a few percent of pixels carry a depth: a
//      thread per pixel left most lanes of a wave idle in the per-view loop.)
//   2. compaction of flagged pixels (order preserving).
//   3. voxel keys floor(p / cell) as i64 and their per-axis range.
//   4. integer voxel hash: 63-bit packed keys (21 bits per axis, relative to
//      the per-axis minimum), open addressing with 64-bit atomicCAS; each
//      candidate learns its slot.
//   5. grouping by slot: member counts per slot (atomics), their exclusive
//      scan, placement of every candidate into its slot's segment, then each
//      segment ordered by candidate index -- candidates of one voxel become
//      contiguous and in reference order, so the f64 coordinate sums and the
//      wrapping u16 colour sums are accumulated in exactly the reference's
//      sequence (fuse.rs:178-201).
//   6. one thread per voxel segment: sums, averages, round() as u8.
//   7. compaction of voxel results by their first candidate index: output
//      in first-insertion order.
// The compactions and scans are this file's own three-phase kernels (block
// sums, a one-block scan of the block sums, block-local scan + scatter).
#include <algorithm>

#include "stereo_kernels.h"

namespace mvs {

struct DevPoint {
    double x, y, z;
    uint8_t r, g, b, pad[5];
};

static constexpr u64 EMPTY_KEY = ~0ULL;

// fuse.rs:32-62
__device__ __forceinline__ double sample_depth(const float* __restrict__ depth, int w, int h, double u, double v) {
    const int32_t u0 = f64_to_i32_sat(floor(u));
    const int32_t v0 = f64_to_i32_sat(floor(v));
    if (u0 < 0 || v0 < 0 || (long long)u0 + 1 >= w || (long long)v0 + 1 >= h) return 0.0;
    const float fx = (float)(u - (double)u0);
    const float fy = (float)(v - (double)v0);
    const size_t r0 = (size_t)v0 * w + u0, r1 = (size_t)(v0 + 1) * w + u0;
    const float d00 = depth[r0], d10 = depth[r0 + 1], d01 = depth[r1], d11 = depth[r1 + 1];
    const float w00 = (1.0f - fx) * (1.0f - fy), w10 = fx * (1.0f - fy), w01 = (1.0f - fx) * fy, w11 = fx * fy;
    float acc = 0.0f, wsum = 0.0f;
    if (d00 > 0.0f) { acc += d00 * w00; wsum += w00; }
    if (d10 > 0.0f) { acc += d10 * w10; wsum += w10; }
    if (d01 > 0.0f) { acc += d01 * w01; wsum += w01; }
    if (d11 > 0.0f) { acc += d11 * w11; wsum += w11; }
    return wsum > 1e-6f ? (double)(acc / wsum) : 0.0;
}

// Pixels with a depth (fuse.rs:98-100 skips the others): vflag[g] = depth > 0
// for every pixel g of every view with a depth map (grid row y = view y),
// and the candidate flag cleared.
__global__ __launch_bounds__(256) void valid_kernel(const FuseView* __restrict__ V, uint8_t* __restrict__ vflag,
                                                    uint8_t* __restrict__ flag) {
    const FuseView& v = V[blockIdx.y];
    if (!v.depth) return;
    const int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= (int64_t)v.w * v.h) return;
    const int64_t g = v.pix_off + li;
    vflag[g] = v.depth[li] > 0.0f ? 1 : 0;
    flag[g] = 0;
}

// fuse.rs:98-163 for one pixel with a depth: thread i takes the i-th such
// pixel (vid[i], ascending), so every lane of a wave runs the per-view loop.
__global__ __launch_bounds__(256) void cand_kernel(const FuseView* __restrict__ V, int nv,
                                                   const int32_t* __restrict__ order, const int32_t* __restrict__ vid,
                                                   int64_t n_valid, int min_cons, double depth_diff,
                                                   uint8_t* __restrict__ flag, double* __restrict__ xyz,
                                                   uchar4* __restrict__ rgb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_valid) return;
    const int64_t g = vid[i];
    int lo = 0, hi = nv - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (V[mid].pix_off <= g) lo = mid; else hi = mid - 1;
    }
    const FuseView& v = V[lo];
    const int64_t li = g - v.pix_off;
    const int px = (int)(li % v.w), py = (int)(li / v.w);
    const double zr = (double)v.depth[li];
    {
        double nx, ny;
        cam_normalize(v.f, v.cx, v.cy, v.k, (double)px, (double)py, nx, ny);
        const double dr[3] = {nx * zr - v.t[0], ny * zr - v.t[1], 1.0 * zr - v.t[2]};
        double xw[3];
        mat_t_vec(v.R, dr, xw);
        double acc[3] = {xw[0], xw[1], xw[2]};
        int ncons = 0;
        for (int q = 0; q < v.ord_n; ++q) {
            const FuseView& s = V[order[v.ord_off + q]];
            if (!s.depth) continue;
            double xs[3];
            mat_vec(s.R, xw, xs);
            xs[0] = xs[0] + s.t[0];
            xs[1] = xs[1] + s.t[1];
            xs[2] = xs[2] + s.t[2];
            if (xs[2] <= 1e-6) continue;
            // Pose::project (geom.rs:63-76): the same x_cam, so z > 1e-9 holds
            const double xn = xs[0] / xs[2], yn = xs[1] / xs[2];
            const double r2 = xn * xn + yn * yn;
            const double dd = 1.0 + s.k * r2;
            const double us = s.f * xn * dd + s.cx, vs = s.f * yn * dd + s.cy;
            const double zs = sample_depth(s.depth, s.w, s.h, us, vs);
            if (zs <= 0.0) continue;
            const double rel = fabs((xs[2] - zs) / xs[2]);
            if (rel < depth_diff) {
                double sx, sy;
                cam_normalize(s.f, s.cx, s.cy, s.k, us, vs, sx, sy);
                const double d2[3] = {sx * zs - s.t[0], sy * zs - s.t[1], 1.0 * zs - s.t[2]};
                double xw2[3];
                mat_t_vec(s.R, d2, xw2);
                acc[0] = acc[0] + xw2[0];
                acc[1] = acc[1] + xw2[1];
                acc[2] = acc[2] + xw2[2];
                ++ncons;
            }
        }
        if (ncons >= min_cons) {
            const double m = (double)(ncons + 1);
            xyz[3 * g] = acc[0] / m;
            xyz[3 * g + 1] = acc[1] / m;
            xyz[3 * g + 2] = acc[2] / m;
            uchar4 c;
            if (v.rgb) {
                c.x = v.rgb[3 * li];
                c.y = v.rgb[3 * li + 1];
                c.z = v.rgb[3 * li + 2];
            } else {
                c.x = c.y = c.z = v.gray[li];
            }
            c.w = 0;
            rgb[g] = c;
            flag[g] = 1;
        }
    }
}

__global__ void key_init_kernel(long long* kr) {
    if (threadIdx.x < 3) kr[threadIdx.x] = 9223372036854775807LL;
    else if (threadIdx.x < 6) kr[threadIdx.x] = (-9223372036854775807LL - 1);
}

// fuse.rs:180-184: floor(p / cell) as i64, and the per-axis key range: each
// thread takes KEY_PER_THREAD candidates (grid stride), then the range is
// reduced over the wave and the block and one atomic per block and axis goes
// to memory (one per wave and axis made ~20 k waves serialise on six words).
constexpr int KEY_PER_THREAD = 8;

__global__ __launch_bounds__(256) void key_kernel(const int32_t* __restrict__ cid, int n, const double* __restrict__ xyz,
                                                  double cell, long long* __restrict__ key3, long long* __restrict__ kr) {
    __shared__ long long s_mn[3][4], s_mx[3][4];
    long long mn[3], mx[3];
    for (int a = 0; a < 3; ++a) {
        mn[a] = 9223372036854775807LL;
        mx[a] = (-9223372036854775807LL - 1);
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int64_t g = cid[i];
        for (int a = 0; a < 3; ++a) {
            const long long k = f64_to_i64_sat(floor(xyz[3 * g + a] / cell));
            key3[3 * (size_t)i + a] = k;
            mn[a] = k < mn[a] ? k : mn[a];
            mx[a] = k > mx[a] ? k : mx[a];
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int a = 0; a < 3; ++a) {
        for (int off = 32; off > 0; off >>= 1) {
            const long long omn = __shfl_xor(mn[a], off, 64);
            const long long omx = __shfl_xor(mx[a], off, 64);
            mn[a] = omn < mn[a] ? omn : mn[a];
            mx[a] = omx > mx[a] ? omx : mx[a];
        }
        if (lane == 0) {
            s_mn[a][wave] = mn[a];
            s_mx[a][wave] = mx[a];
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int a = threadIdx.x;
        long long bmn = s_mn[a][0], bmx = s_mx[a][0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
            bmn = s_mn[a][w] < bmn ? s_mn[a][w] : bmn;
            bmx = s_mx[a][w] > bmx ? s_mx[a][w] : bmx;
        }
        atomicMin(&kr[a], bmn);
        atomicMax(&kr[3 + a], bmx);
    }
}

__device__ __forceinline__ u64 mix64(u64 k) {
    k ^= k >> 33;
    k *= 0xFF51AFD7ED558CCDULL;
    k ^= k >> 33;
    k *= 0xC4CEB9FE1A85EC53ULL;
    k ^= k >> 33;
    return k;
}

__global__ __launch_bounds__(256) void insert_kernel(const long long* __restrict__ key3, int n,
                                                     const long long* __restrict__ kr, u64* __restrict__ table,
                                                     u64 cap_mask, uint32_t* __restrict__ slot_of,
                                                     int* __restrict__ err) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u64 key = ((u64)(key3[3 * (size_t)i] - kr[0]) << 42) | ((u64)(key3[3 * (size_t)i + 1] - kr[1]) << 21) |
                    (u64)(key3[3 * (size_t)i + 2] - kr[2]);
    u64 h = mix64(key) & cap_mask;
    for (u64 probe = 0; probe <= cap_mask; ++probe) {
        const u64 prev = atomicCAS((unsigned long long*)&table[h], (unsigned long long)EMPTY_KEY,
                                   (unsigned long long)key);
        if (prev == EMPTY_KEY || prev == key) {
            slot_of[i] = (uint32_t)h;
            return;
        }
        h = (h + 1) & cap_mask;
    }
    atomicExch(err, 1);   // table full: impossible with cap >= 2n, reported anyway
    slot_of[i] = 0;
}

__global__ void fill_u32_k(uint32_t* p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void fill_u64_k(u64* p, size_t n, u64 v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void fill_u8_k(uint8_t* p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// --------------------------------------------------------------------------
// order-preserving compaction and exclusive scan (three phases)
// --------------------------------------------------------------------------

static constexpr int SCAN_T = 256;             // threads per block of phases 1 and 3
static constexpr int SCAN_I = 16;              // consecutive items per thread
static constexpr int SCAN_B = SCAN_T * SCAN_I; // items per block
static constexpr int SCAN2_T = 1024;           // threads of the one-block phase 2

// exclusive scan of one value per thread over a block of NT threads; returns
// the block total
template <int NT>
__device__ __forceinline__ uint32_t block_exscan(uint32_t v, uint32_t& total) {
    __shared__ uint32_t s_w[NT / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_w[wv] = x;
    __syncthreads();
    uint32_t base = 0, tot = 0;
    for (int q = 0; q < NT / 64; ++q) {
        if (q < wv) base += s_w[q];
        tot += s_w[q];
    }
    __syncthreads();   // s_w is reused by the next call
    total = tot;
    return base + x - v;
}

// phase 1: the number of selected items (flag != 0) / the sum of counts of each block
template <class T>
__global__ __launch_bounds__(SCAN_T) void block_sum_kernel(const T* __restrict__ in, int64_t n,
                                                           uint32_t* __restrict__ bsum) {
    const int64_t b0 = (int64_t)blockIdx.x * SCAN_B + (int64_t)threadIdx.x * SCAN_I;
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < SCAN_I; ++k)
        if (b0 + k < n) c += (uint32_t)in[b0 + k];
    uint32_t tot;
    (void)block_exscan<SCAN_T>(c, tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// phase 2: exclusive scan of the nb block sums in place (one block); total -> *total
__global__ __launch_bounds__(SCAN2_T) void scan_bsum_kernel(uint32_t* __restrict__ bsum, int nb,
                                                            uint32_t* __restrict__ total) {
    uint32_t carry = 0;
    for (int base = 0; base < nb; base += SCAN2_T) {
        const int i = base + (int)threadIdx.x;
        const uint32_t v = i < nb ? bsum[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exscan<SCAN2_T>(v, tot);
        if (i < nb) bsum[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

// phase 3 (compaction): selected item g -> out[position] = g (IDX) or src[g]
template <class T, bool IDX>
__global__ __launch_bounds__(SCAN_T) void compact_kernel(const uint8_t* __restrict__ flag, int64_t n,
                                                         const uint32_t* __restrict__ boff, const T* __restrict__ src,
                                                         T* __restrict__ out) {
    const int64_t b0 = (int64_t)blockIdx.x * SCAN_B + (int64_t)threadIdx.x * SCAN_I;
    uint8_t f[SCAN_I];
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < SCAN_I; ++k) {
        f[k] = b0 + k < n ? flag[b0 + k] : 0;
        c += f[k] != 0;
    }
    uint32_t tot;
    uint32_t pos = boff[blockIdx.x] + block_exscan<SCAN_T>(c, tot);
#pragma unroll
    for (int k = 0; k < SCAN_I; ++k)
        if (f[k]) {
            if constexpr (IDX) out[pos] = (T)(b0 + k);
            else out[pos] = src[b0 + k];
            ++pos;
        }
}

// phase 3 (scan): out[i] = out2[i] = exclusive prefix sum of in[] at i
__global__ __launch_bounds__(SCAN_T) void exscan_kernel(const uint32_t* __restrict__ in, int64_t n,
                                                        const uint32_t* __restrict__ boff, uint32_t* __restrict__ out,
                                                        uint32_t* __restrict__ out2) {
    const int64_t b0 = (int64_t)blockIdx.x * SCAN_B + (int64_t)threadIdx.x * SCAN_I;
    uint32_t v[SCAN_I];
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < SCAN_I; ++k) {
        v[k] = b0 + k < n ? in[b0 + k] : 0u;
        c += v[k];
    }
    uint32_t tot;
    uint32_t run = boff[blockIdx.x] + block_exscan<SCAN_T>(c, tot);
#pragma unroll
    for (int k = 0; k < SCAN_I; ++k)
        if (b0 + k < n) {
            out[b0 + k] = run;
            out2[b0 + k] = run;
            run += v[k];
        }
}

// --------------------------------------------------------------------------
// grouping of candidates by hash slot
// --------------------------------------------------------------------------

__global__ __launch_bounds__(256) void slot_count_kernel(const uint32_t* __restrict__ slot_of, int n,
                                                         uint32_t* __restrict__ cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(&cnt[slot_of[i]], 1u);
}

// candidate i into its slot's segment [off[s], off[s] + cnt[s]) (cursor = a
// copy of off); the order inside a segment is fixed afterwards
__global__ __launch_bounds__(256) void slot_place_kernel(const uint32_t* __restrict__ slot_of, int n,
                                                         uint32_t* __restrict__ cursor, uint32_t* __restrict__ member) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) member[atomicAdd(&cursor[slot_of[i]], 1u)] = (uint32_t)i;
}

// One thread per occupied slot (fuse.rs:185-212): its members ordered by
// candidate index (insertion sort in place; a voxel holds few candidates),
// then the f64 sums and wrapping u16 colour sums in that order, the
// averages and round() as u8, stored at the first candidate's index.
__global__ __launch_bounds__(256) void reduce_kernel(const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ off,
                                                     int64_t nslot, uint32_t* __restrict__ member,
                                                     const int32_t* __restrict__ cid, const double* __restrict__ xyz,
                                                     const uchar4* __restrict__ rgb, DevPoint* __restrict__ by_first,
                                                     uint8_t* __restrict__ first_flag) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslot) return;
    const uint32_t m = cnt[s];
    if (m == 0) return;
    uint32_t* seg = member + off[s];
    for (uint32_t a = 1; a < m; ++a) {
        const uint32_t v = seg[a];
        uint32_t b = a;
        for (; b > 0 && seg[b - 1] > v; --b) seg[b] = seg[b - 1];
        seg[b] = v;
    }
    double sx = 0.0, sy = 0.0, sz = 0.0;
    uint16_t r = 0, g = 0, b = 0;
    for (uint32_t q = 0; q < m; ++q) {
        const int64_t gi = cid[seg[q]];
        sx = sx + xyz[3 * gi];
        sy = sy + xyz[3 * gi + 1];
        sz = sz + xyz[3 * gi + 2];
        const uchar4 c = rgb[gi];
        r = (uint16_t)(r + c.x);
        g = (uint16_t)(g + c.y);
        b = (uint16_t)(b + c.z);
    }
    const double md = (double)m;
    DevPoint o;
    o.x = sx / md;
    o.y = sy / md;
    o.z = sz / md;
    o.r = f64_to_u8_sat(round((double)r / md));
    o.g = f64_to_u8_sat(round((double)g / md));
    o.b = f64_to_u8_sat(round((double)b / md));
    for (int k = 0; k < 5; ++k) o.pad[k] = 0;
    const uint32_t first = seg[0];
    by_first[first] = o;
    first_flag[first] = 1;
}

// --------------------------------------------------------------------------
// host driver
// --------------------------------------------------------------------------

template <class T>
struct Buf {
    T* p = nullptr;
    size_t cap = 0;
    hipError_t need(size_t n) {
        if (p && n <= cap) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        const size_t c = n < 1024 ? 1024 : n + n / 8;
        hipError_t e = hipMalloc((void**)&p, c * sizeof(T));
        if (e == hipSuccess) cap = c;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

struct FuseScratch {
    Buf<uint8_t> flag, first_flag, vflag;
    Buf<int32_t> vid;
    Buf<double> xyz;
    Buf<uchar4> rgb;
    Buf<int32_t> cid, counts;   // counts: [0] n_sel, [1] n_pts, [2] err
    Buf<long long> key3, kr;    // kr: per-axis key min (3) and max (3)
    Buf<uint32_t> slot, cnt, off, cursor, member, bsum;
    Buf<DevPoint> by_first, out;
    Buf<u64> table;
};

#define FCHK(x)                                 \
    do {                                        \
        hipError_t e_ = (x);                    \
        if (e_ != hipSuccess) return e_;        \
    } while (0)

void free_fuse_scratch(FuseScratch* s) {
    if (!s) return;
    s->flag.release(); s->first_flag.release(); s->vflag.release(); s->vid.release();
    s->xyz.release(); s->rgb.release();
    s->cid.release(); s->counts.release(); s->key3.release(); s->kr.release(); s->slot.release();
    s->cnt.release(); s->off.release(); s->cursor.release(); s->member.release(); s->bsum.release();
    s->by_first.release(); s->out.release(); s->table.release();
    delete s;
}

static inline unsigned nblk(size_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }

hipError_t run_fusion(FuseScratch*& S, const FuseArgs& a, const FuseView* h_views, hipStream_t st, int64_t* n_cand,
                      int64_t* n_pts, const void** d_points, int* limit_err) {
    *limit_err = 0;
    *n_cand = 0;
    *n_pts = 0;
    if (!S) S = new FuseScratch();
    FCHK(S->kr.need(6));
    FCHK(S->counts.need(4));
    const size_t np = (size_t)a.n_pix;
    if (np == 0) return hipSuccess;
    if (np > 0x7fffffffULL) {
        *limit_err = 1;
        return hipSuccess;
    }
    FCHK(S->flag.need(np));
    FCHK(S->vflag.need(np));
    FCHK(S->vid.need(np));
    FCHK(S->xyz.need(3 * np));
    FCHK(S->rgb.need(np));
    FCHK(S->cid.need(np));
    const size_t nb1 = (np + SCAN_B - 1) / SCAN_B;
    FCHK(S->bsum.need(nb1 + 1));
    uint32_t* total = S->bsum.p + nb1;
    // pixels with a depth, then their list in global order (vid)
    int64_t max_px = 0;
    for (int i = 0; i < a.nviews; ++i)
        if (h_views[i].depth) max_px = std::max(max_px, (int64_t)h_views[i].w * h_views[i].h);
    hipLaunchKernelGGL(valid_kernel, dim3(nblk((size_t)max_px), (unsigned)a.nviews), dim3(256), 0, st, a.d_views,
                       S->vflag.p, S->flag.p);
    hipLaunchKernelGGL((block_sum_kernel<uint8_t>), dim3((unsigned)nb1), dim3(SCAN_T), 0, st, S->vflag.p,
                       (int64_t)np, S->bsum.p);
    hipLaunchKernelGGL(scan_bsum_kernel, dim3(1), dim3(SCAN2_T), 0, st, S->bsum.p, (int)nb1, total);
    hipLaunchKernelGGL((compact_kernel<int32_t, true>), dim3((unsigned)nb1), dim3(SCAN_T), 0, st, S->vflag.p,
                       (int64_t)np, S->bsum.p, (const int32_t*)nullptr, S->vid.p);
    FCHK(hipGetLastError());
    uint32_t nvalid = 0;
    FCHK(hipMemcpyAsync(&nvalid, total, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    FCHK(hipStreamSynchronize(st));
    if (nvalid > 0)
        hipLaunchKernelGGL(cand_kernel, dim3(nblk(nvalid)), dim3(256), 0, st, a.d_views, a.nviews, a.d_order,
                           S->vid.p, (int64_t)nvalid, a.min_consistent, a.depth_diff, S->flag.p, S->xyz.p, S->rgb.p);
    FCHK(hipGetLastError());
    // compaction of candidates (order preserving): cid[k] = index of the k-th flagged pixel
    hipLaunchKernelGGL((block_sum_kernel<uint8_t>), dim3((unsigned)nb1), dim3(SCAN_T), 0, st, S->flag.p, (int64_t)np,
                       S->bsum.p);
    hipLaunchKernelGGL(scan_bsum_kernel, dim3(1), dim3(SCAN2_T), 0, st, S->bsum.p, (int)nb1, total);
    hipLaunchKernelGGL((compact_kernel<int32_t, true>), dim3((unsigned)nb1), dim3(SCAN_T), 0, st, S->flag.p,
                       (int64_t)np, S->bsum.p, (const int32_t*)nullptr, S->cid.p);
    FCHK(hipGetLastError());
    uint32_t nsel = 0;
    FCHK(hipMemcpyAsync(&nsel, total, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    FCHK(hipStreamSynchronize(st));
    *n_cand = nsel;
    if (nsel == 0) return hipSuccess;
    const size_t n = (size_t)nsel;
    FCHK(S->key3.need(3 * n));
    FCHK(S->slot.need(n));
    FCHK(S->member.need(n));
    FCHK(S->by_first.need(n));
    FCHK(S->first_flag.need(n));
    FCHK(S->out.need(n));
    hipLaunchKernelGGL(key_init_kernel, dim3(1), dim3(64), 0, st, S->kr.p);
    hipLaunchKernelGGL(key_kernel, dim3(nblk((n + KEY_PER_THREAD - 1) / KEY_PER_THREAD)), dim3(256), 0, st, S->cid.p,
                       (int)n, S->xyz.p, a.cell, S->key3.p, S->kr.p);
    FCHK(hipGetLastError());
    long long kr[6];
    FCHK(hipMemcpyAsync(kr, S->kr.p, sizeof(kr), hipMemcpyDeviceToHost, st));
    FCHK(hipStreamSynchronize(st));
    for (int ax = 0; ax < 3; ++ax) {
        // the packed key holds 21 bits per axis (2,097,151 cells of extent)
        const unsigned long long ext = (unsigned long long)kr[3 + ax] - (unsigned long long)kr[ax];
        if (ext >= (1ULL << 21)) {
            *limit_err = 2;
            return hipSuccess;
        }
    }
    size_t cap = 1024;
    while (cap < 2 * n) cap <<= 1;
    FCHK(S->table.need(cap));
    FCHK(S->cnt.need(cap));
    FCHK(S->off.need(cap));
    FCHK(S->cursor.need(cap));
    hipLaunchKernelGGL(fill_u64_k, dim3(nblk(cap)), dim3(256), 0, st, S->table.p, cap, EMPTY_KEY);
    hipLaunchKernelGGL(fill_u64_k, dim3(1), dim3(64), 0, st, (u64*)S->counts.p, (size_t)2, 0ULL);
    hipLaunchKernelGGL(fill_u32_k, dim3(nblk(cap)), dim3(256), 0, st, S->cnt.p, cap);
    hipLaunchKernelGGL(insert_kernel, dim3(nblk(n)), dim3(256), 0, st, S->key3.p, (int)n, S->kr.p, S->table.p,
                       (u64)(cap - 1), S->slot.p, S->counts.p + 2);
    // grouping by slot: counts, their exclusive scan (segment offsets), placement
    hipLaunchKernelGGL(slot_count_kernel, dim3(nblk(n)), dim3(256), 0, st, S->slot.p, (int)n, S->cnt.p);
    const size_t nb2 = (cap + SCAN_B - 1) / SCAN_B;
    FCHK(S->bsum.need(nb1 > nb2 ? nb1 + 1 : nb2 + 1));
    total = S->bsum.p + nb2;
    hipLaunchKernelGGL((block_sum_kernel<uint32_t>), dim3((unsigned)nb2), dim3(SCAN_T), 0, st, S->cnt.p,
                       (int64_t)cap, S->bsum.p);
    hipLaunchKernelGGL(scan_bsum_kernel, dim3(1), dim3(SCAN2_T), 0, st, S->bsum.p, (int)nb2, total);
    hipLaunchKernelGGL(exscan_kernel, dim3((unsigned)nb2), dim3(SCAN_T), 0, st, S->cnt.p, (int64_t)cap, S->bsum.p,
                       S->off.p, S->cursor.p);
    hipLaunchKernelGGL(slot_place_kernel, dim3(nblk(n)), dim3(256), 0, st, S->slot.p, (int)n, S->cursor.p,
                       S->member.p);
    hipLaunchKernelGGL(fill_u8_k, dim3(nblk(n)), dim3(256), 0, st, S->first_flag.p, n);
    hipLaunchKernelGGL(reduce_kernel, dim3(nblk(cap)), dim3(256), 0, st, S->cnt.p, S->off.p, (int64_t)cap,
                       S->member.p, S->cid.p, S->xyz.p, S->rgb.p, S->by_first.p, S->first_flag.p);
    FCHK(hipGetLastError());
    // voxel results in first-insertion order
    const size_t nb3 = (n + SCAN_B - 1) / SCAN_B;
    total = S->bsum.p + nb3;
    hipLaunchKernelGGL((block_sum_kernel<uint8_t>), dim3((unsigned)nb3), dim3(SCAN_T), 0, st, S->first_flag.p,
                       (int64_t)n, S->bsum.p);
    hipLaunchKernelGGL(scan_bsum_kernel, dim3(1), dim3(SCAN2_T), 0, st, S->bsum.p, (int)nb3, total);
    hipLaunchKernelGGL((compact_kernel<DevPoint, false>), dim3((unsigned)nb3), dim3(SCAN_T), 0, st, S->first_flag.p,
                       (int64_t)n, S->bsum.p, S->by_first.p, S->out.p);
    FCHK(hipGetLastError());
    uint32_t npts = 0;
    int cnt[3] = {0, 0, 0};
    FCHK(hipMemcpyAsync(&npts, total, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    FCHK(hipMemcpyAsync(cnt, S->counts.p, sizeof(cnt), hipMemcpyDeviceToHost, st));
    FCHK(hipStreamSynchronize(st));
    if (cnt[2]) {
        *limit_err = 3;
        return hipSuccess;
    }
    cnt[1] = (int)npts;
    *n_pts = cnt[1];
    *d_points = S->out.p;
    return hipSuccess;
}

}  // namespace mvs
