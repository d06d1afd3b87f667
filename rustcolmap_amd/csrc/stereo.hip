// stereo.hip -- census, ray setup, plane-sweep init, propagation passes and
// validity gate of the dense stereo path, written for gfx950 (wave64).
//
// Reference semantics: mvs/src/lib.rs:119-142 (census_map) and
// mvs/src/stereo.rs:114-394 (RefWorkspace::new, plane_cost, estimate_depth).
//
// Work decomposition: every stage is one launch over a tile space that
// concatenates the tiles of all jobs (reference views) of a batch, so views
// of uneven size and support-set size K share a launch.  A workgroup owns
// one TW x TH pixel tile of one job; the reference census and rays of the
// tile plus a 2-pixel apron (the 5x5 cost window) are staged in LDS with
// clamp-to-edge addressing, so every window tap of every candidate plane is
// an LDS read; source-census taps are plane-dependent gathers served by
// L1/L2.  A pass reads plane buffer `par` and writes `par ^ 1` for every
// pixel (Jacobi, exactly the reference's double-buffered pass,
// stereo.rs:271-371).
#include "mvs_device.h"
#include "stereo_kernels.h"

namespace mvs {

// --------------------------------------------------------------------------
// census + rays for all views of an upload, one launch
// --------------------------------------------------------------------------

constexpr int CT_W = 32, CT_H = 8;

__global__ __launch_bounds__(CT_W* CT_H) void prepare_kernel(const PrepView* __restrict__ views, int nviews) {
    // view of this block: last view whose tile_off <= blockIdx.x (uniform)
    int lo = 0, hi = nviews - 1;
    const int b = blockIdx.x;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (views[mid].tile_off <= b) lo = mid; else hi = mid - 1;
    }
    const PrepView& V = views[lo];
    const int w = V.w, h = V.h;
    const int t = b - V.tile_off;
    const int x0 = (t % V.tiles_x) * CT_W, y0 = (t / V.tiles_x) * CT_H;
    __shared__ uint8_t s[CT_H + 6][CT_W + 6];
    for (int i = threadIdx.x; i < (CT_H + 6) * (CT_W + 6); i += CT_W * CT_H) {
        const int ay = i / (CT_W + 6), ax = i % (CT_W + 6);
        const int gx = min(max(x0 - 3 + ax, 0), w - 1);
        const int gy = min(max(y0 - 3 + ay, 0), h - 1);
        s[ay][ax] = V.gray[(size_t)gy * w + gx];
    }
    __syncthreads();
    const int lx = threadIdx.x % CT_W, ly = threadIdx.x / CT_W;
    const int x = x0 + lx, y = y0 + ly;
    if (x >= w || y >= h) return;
    const uint8_t c = s[ly + 3][lx + 3];
    u64 bits = 0;
#pragma unroll
    for (int dy = -3; dy <= 3; dy++)
#pragma unroll
        for (int dx = -3; dx <= 3; dx++) {
            if (dx == 0 && dy == 0) continue;
            if (s[ly + 3 + dy][lx + 3 + dx] < c) bits |= 1ULL << ((dy + 3) * 7 + (dx + 3));
        }
    const size_t i = (size_t)y * w + x;
    if (V.census) V.census[i] = bits;
    if (V.rays) {
        double xo, yo;
        cam_normalize(V.f, V.cx, V.cy, V.k, (double)x, (double)y, xo, yo);
        V.rays[i] = make_float2((float)xo, (float)yo);
    }
}

// --------------------------------------------------------------------------
// plane_cost (stereo.rs:158-228)
// --------------------------------------------------------------------------

// Reference-window access: LDS tile with apron (production) ...
struct LdsRef {
    const u64* cen;
    const float2* ray;
    int c, pitch;
    __device__ __forceinline__ int idx(int dy, int dx) const { return c + dy * pitch + dx; }
};
// ... or clamped global reads (instrumentation kernel).
struct GlobRef {
    const u64* cen;
    const float2* ray;
    int px, py, w, h;
    __device__ __forceinline__ int idx(int dy, int dx) const {
        const int yy = min(max(py + dy, 0), h - 1), xx = min(max(px + dx, 0), w - 1);
        return yy * w + xx;
    }
};

struct EvalInfo {
    int ham, ninv;
    float peak;
};

// Cost of plane (n, d) at the window's centre pixel.  `thr` is
// (hint + 0.12) with hint = best + 0.02.  Abort iff some (source, row)
// prefix average exceeds thr; with INFO the full window is evaluated and
// the abort is decided from the peak prefix average, which is the same
// predicate.
template <bool INFO, class Ref>
__device__ __forceinline__ float eval_plane(const Ref& R, const SrcDesc* __restrict__ S, int nsrc, float nx,
                                            float ny, float nz, float d, float thr, EvalInfo* info) {
    float sum = 0.0f;
    int ham = 0, ninv = 0;
    float peak = -__builtin_inff();
    bool aborted = false;
#pragma nounroll
    for (int si = 0; si < nsrc; ++si) {
        const SrcDesc& s = S[si];
        const u64* __restrict__ cen = s.census;
        const int sw = s.w, sh = s.h;
        const float r0 = s.rs[0], r1 = s.rs[1], r2 = s.rs[2], r3 = s.rs[3], r4 = s.rs[4], r5 = s.rs[5];
        const float r6 = s.rs[6], r7 = s.rs[7], r8 = s.rs[8];
        const float t0 = s.ts[0], t1 = s.ts[1], t2 = s.ts[2];
        const float f = s.f, cx = s.cx, cy = s.cy, kk = s.k;
#pragma unroll
        for (int dy = -2; dy <= 2; ++dy) {
#pragma unroll
            for (int dx = -2; dx <= 2; ++dx) {
                const int j = R.idx(dy, dx);
                const float2 r = R.ray[j];
                const float denom = (nx * r.x + ny * r.y) + nz;
                if (denom < 0.05f) { sum += 1.0f; ++ninv; continue; }
                const float z = d / denom;
                const float xr = z * r.x, yr = z * r.y;
                const float xs = ((r0 * xr + r1 * yr) + r2 * z) + t0;
                const float ys = ((r3 * xr + r4 * yr) + r5 * z) + t1;
                const float zs = ((r6 * xr + r7 * yr) + r8 * z) + t2;
                if (zs < 1e-3f) { sum += 1.0f; ++ninv; continue; }
                const float xn = xs / zs, yn = ys / zs;
                const float dist = 1.0f + kk * (xn * xn + yn * yn);
                int u, v;
                const bool oku = cast_in_range(f * xn * dist + cx, sw, u);
                const bool okv = cast_in_range(f * yn * dist + cy, sh, v);
                if (!(oku && okv)) { sum += 1.0f; ++ninv; continue; }
                const int pc = __popcll(R.cen[j] ^ cen[(size_t)v * sw + u]);
                ham += pc;
                sum += pc_over_48(pc);
            }
            const float part = sum / ((float)(dy + 3) * 5.0f * (float)(si + 1));
            if (INFO) {
                peak = fmaxf(peak, part);
                aborted = aborted || (part > thr);
            } else if (part > thr) {
                return 1e9f;
            }
        }
    }
    if (INFO) {
        info->ham = ham;
        info->ninv = ninv;
        info->peak = peak;
        if (aborted) return 1e9f;
    }
    return sum / (25.0f * (float)nsrc);
}

// Production evaluator.  Same value as eval_plane<false>, organised for the
// GPU:
//   * the window geometry (denom test, z = d / denom, z*xo, z*yo) does not
//     depend on the source, so it is computed once, while source 0 is
//     evaluated, and kept in registers for the other sources -- the same
//     operations on the same operands, hence the same values;
//   * taps are branch-free: an invalid tap still issues a (clamped) gather
//     and adds 1.0 by select, so the five gathers of a window row are in
//     flight together;
//   * source census reads go through the global address space.
typedef __attribute__((address_space(1))) const u64 gu64;
// global-address-space views of a plane (as a plain vector type: float4's
// class assignment does not take another address space) and of a cost
typedef float f32x4v __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const f32x4v gplane;
typedef __attribute__((address_space(1))) const float gfloat;
__device__ __forceinline__ float4 as_float4(f32x4v v) { return make_float4(v.x, v.y, v.z, v.w); }

struct SrcP {
    gu64* cen;
    __amdgpu_buffer_rsrc_t rcen;   // the same census as a raw buffer (scalar registers)
    int sw, sh;
    float r0, r1, r2, r3, r4, r5, r6, r7, r8, t0, t1, t2, f, cx, cy, k;
};

// gfx9 raw-buffer descriptor word 3 (32-bit data, no swizzle)
constexpr int RAW_BUFFER_FLAGS = 0x00020000;

__device__ __forceinline__ SrcP load_src(const SrcDesc& s) {
    SrcP p;
    p.cen = (gu64*)s.census;
    // record count in bytes: a view holds at most 2^29 pixels (host check)
    const unsigned long long bytes = (unsigned long long)s.w * (unsigned long long)s.h * 8ull;
    p.rcen = __builtin_amdgcn_make_buffer_rsrc((void*)s.census, (short)0,
                                               (int)(bytes > 0xffffffffull ? 0xffffffffull : bytes), RAW_BUFFER_FLAGS);
    p.sw = s.w;
    p.sh = s.h;
    p.r0 = s.rs[0]; p.r1 = s.rs[1]; p.r2 = s.rs[2]; p.r3 = s.rs[3]; p.r4 = s.rs[4];
    p.r5 = s.rs[5]; p.r6 = s.rs[6]; p.r7 = s.rs[7]; p.r8 = s.rs[8];
    p.t0 = s.ts[0]; p.t1 = s.ts[1]; p.t2 = s.ts[2];
    p.f = s.f; p.cx = s.cx; p.cy = s.cy; p.k = s.k;
    return p;
}

// `(x) as i32` lands in [0, n): NaN -> 0 (valid for n > 0), else -1 < x < n.
// The census word at byte offset `off` (< the census size: callers pass 0
// for an invalid tap) through a raw-buffer load: the view's base stays in
// scalar registers and the offset is a 32-bit vector register, so a tap has
// no 64-bit address arithmetic.
__device__ __forceinline__ u64 cen_at(const SrcP& P, unsigned off) {
    return __builtin_bit_cast(u64, __builtin_amdgcn_raw_buffer_load_b64(P.rcen, (int)off, 0, 0));
}

__device__ __forceinline__ bool cast_ok(float x, int n, int& out) {
    const bool in = (x > -1.0f) && (x < (float)n);
    out = (int)(in ? x : 0.0f);
    return in || (x != x);
}

// One (window tap, source) of stereo.rs:187-213 given the window geometry.
__device__ __forceinline__ bool tap_index(float z, float xr, float yr, bool vd, const SrcP& P, int& idx) {
    const float xs = ((P.r0 * xr + P.r1 * yr) + P.r2 * z) + P.t0;
    const float ys = ((P.r3 * xr + P.r4 * yr) + P.r5 * z) + P.t1;
    const float zs = ((P.r6 * xr + P.r7 * yr) + P.r8 * z) + P.t2;
    const float xn = xs / zs, yn = ys / zs;
    const float dist = 1.0f + P.k * (xn * xn + yn * yn);
    int u, v;
    const bool oku = cast_ok(P.f * xn * dist + P.cx, P.sw, u);
    const bool okv = cast_ok(P.f * yn * dist + P.cy, P.sh, v);
    const bool ok = vd && !(zs < 1e-3f) && oku && okv;
    idx = ok ? v * P.sw + u : 0;
    return ok;
}

template <int PW>
__device__ __forceinline__ float eval_plane_fast(const u64* __restrict__ s_cen, const float2* __restrict__ s_ray, int c,
                                                 const SrcDesc* __restrict__ S, int nsrc, float nx, float ny, float nz,
                                                 float d, float thr) {
    float gz[25], gx[25], gy[25];
    unsigned okd = 0;
    float sum = 0.0f;
    {
        const SrcP P = load_src(S[0]);
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            int idx[5];
            bool ok[5];
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                const int t = r * 5 + q;
                const float2 ray = s_ray[c + (r - 2) * PW + (q - 2)];
                const float denom = (nx * ray.x + ny * ray.y) + nz;
                const bool vd = !(denom < 0.05f);
                const float z = d / denom;
                gz[t] = z;
                gx[t] = z * ray.x;
                gy[t] = z * ray.y;
                okd |= (vd ? 1u : 0u) << t;
                ok[q] = tap_index(gz[t], gx[t], gy[t], vd, P, idx[q]);
            }
            u64 cs[5];
#pragma unroll
            for (int q = 0; q < 5; ++q) cs[q] = P.cen[idx[q]];
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                const u64 rc = s_cen[c + (r - 2) * PW + (q - 2)];
                sum += ok[q] ? pc_over_48(__popcll(rc ^ cs[q])) : 1.0f;
            }
            if (sum / ((float)(r + 1) * 5.0f) > thr) return 1e9f;
        }
    }
#pragma nounroll
    for (int si = 1; si < nsrc; ++si) {
        const SrcP P = load_src(S[si]);
        const float fs = (float)(si + 1);
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            int idx[5];
            bool ok[5];
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                const int t = r * 5 + q;
                ok[q] = tap_index(gz[t], gx[t], gy[t], (okd >> t) & 1u, P, idx[q]);
            }
            u64 cs[5];
#pragma unroll
            for (int q = 0; q < 5; ++q) cs[q] = P.cen[idx[q]];
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                const u64 rc = s_cen[c + (r - 2) * PW + (q - 2)];
                sum += ok[q] ? pc_over_48(__popcll(rc ^ cs[q])) : 1.0f;
            }
            if (sum / ((float)(r + 1) * 5.0f * fs) > thr) return 1e9f;
        }
    }
    return sum / (25.0f * (float)nsrc);
}

// Third evaluator, same values again, fewer VALU operations per tap:
//   * the per-row abort test RN(sum / D) > thr is decided without a divide:
//     with m the midpoint of thr and its successor (exact in f64, 25
//     significant bits) the quotient rounds above thr iff sum > m * D, or
//     sum == m * D and thr's last mantissa bit is odd (ties to even); m * D
//     is exact in f64 because D <= 5 * 5 * K has few bits, and the tie
//     cannot occur (make_abort_thr), so a row test is one f64 compare;
//   * sources with k == 0 skip the distortion polynomial: dist = 1 + 0 * r2
//     is exactly 1 whenever r2 is finite, which the host guarantees through
//     a bound on zmax and the translation (JobDesc.pinhole_ok);
//   * `(x) as i32` is v_cvt_i32_f32, which truncates, saturates and maps NaN
//     to 0 exactly like Rust, so the bounds test is one unsigned compare;
//   * popcount / 48 comes from a 49-entry LDS table of the exact quotients.
__device__ __forceinline__ int cvt_i32_rust(float x) {
    int r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

struct AbortThr {
    double m;
};

// RN(sum / D) > thr  <=>  sum > m D, or sum == m D with thr's last mantissa
// bit odd, for m the f64 midpoint of thr and its f32 successor.  The tie
// cannot happen: thr >= 0.02 is a normal float, thr = a 2^e with a 24-bit a,
// so m = (2a + 1) 2^(e-1) has an odd 25-bit significand, and m D (exact in
// f64: D = 5 dv < 2^28) has an odd part of at least 25 bits -- never an f32
// value, which has at most 24.  So the row test is the single compare
// sum > m D.  thr = inf gives m = NaN, and no test fires.
__device__ __forceinline__ AbortThr make_abort_thr(float thr) {
    AbortThr a;
    a.m = 0.5 * ((double)thr + (double)__uint_as_float(__float_as_uint(thr) + 1u));
    return a;
}

// RN(sum / D) > thr for sum >= 0, thr >= 0.02 finite (or never when thr = inf)
__device__ __forceinline__ bool abort_test(float sum, double D, const AbortThr& a) {
    return (double)sum > a.m * D;
}

template <bool PIN>
__device__ __forceinline__ bool tap_index_v3(float z, float xr, float yr, bool vd, const SrcP& P, int& idx) {
    const float xs = ((P.r0 * xr + P.r1 * yr) + P.r2 * z) + P.t0;
    const float ys = ((P.r3 * xr + P.r4 * yr) + P.r5 * z) + P.t1;
    const float zs = ((P.r6 * xr + P.r7 * yr) + P.r8 * z) + P.t2;
    const float xn = xs / zs, yn = ys / zs;
    float fu, fv;
    if (PIN) {
        fu = P.f * xn + P.cx;
        fv = P.f * yn + P.cy;
    } else {
        const float dist = 1.0f + P.k * (xn * xn + yn * yn);
        fu = P.f * xn * dist + P.cx;
        fv = P.f * yn * dist + P.cy;
    }
    const int u = cvt_i32_rust(fu), v = cvt_i32_rust(fv);
    const bool ok = vd && !(zs < 1e-3f) && ((unsigned)u < (unsigned)P.sw) && ((unsigned)v < (unsigned)P.sh);
    idx = ok ? v * P.sw + u : 0;
    return ok;
}

template <int PW, bool PIN>
__device__ __forceinline__ bool row_v3(const u64* __restrict__ s_cen, int c, int r, const float* gz, const float* gx,
                                       const float* gy, unsigned okd, const SrcP& P, const float* __restrict__ tab,
                                       float& sum) {
    int idx[5];
    bool ok[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        const int t = r * 5 + q;
        ok[q] = tap_index_v3<PIN>(gz[t], gx[t], gy[t], (okd >> t) & 1u, P, idx[q]);
    }
    u64 cs[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) cs[q] = P.cen[idx[q]];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        const u64 rc = s_cen[c + (r - 2) * PW + (q - 2)];
        sum += ok[q] ? tab[__popcll(rc ^ cs[q])] : 1.0f;
    }
    return true;
}

template <int PW>
__device__ __forceinline__ float eval_plane_v3(const u64* __restrict__ s_cen, const float2* __restrict__ s_ray, int c,
                                               const SrcDesc* __restrict__ S, int nsrc, float nx, float ny, float nz,
                                               float d, float thr, const float* __restrict__ tab, bool pin_ok) {
    float gz[25], gx[25], gy[25];
    unsigned okd = 0;
#pragma unroll
    for (int t = 0; t < 25; ++t) {
        const float2 ray = s_ray[c + (t / 5 - 2) * PW + (t % 5 - 2)];
        const float denom = (nx * ray.x + ny * ray.y) + nz;
        okd |= (!(denom < 0.05f) ? 1u : 0u) << t;
        const float z = d / denom;
        gz[t] = z;
        gx[t] = z * ray.x;
        gy[t] = z * ray.y;
    }
    const AbortThr A = make_abort_thr(thr);
    float sum = 0.0f;
#pragma nounroll
    for (int si = 0; si < nsrc; ++si) {
        const SrcP P = load_src(S[si]);
        const double fs = (double)(si + 1) * 5.0;
        if (pin_ok && P.k == 0.0f) {
#pragma unroll
            for (int r = 0; r < 5; ++r) {
                row_v3<PW, true>(s_cen, c, r, gz, gx, gy, okd, P, tab, sum);
                if (abort_test(sum, (double)(r + 1) * fs, A)) return 1e9f;
            }
        } else {
#pragma unroll
            for (int r = 0; r < 5; ++r) {
                row_v3<PW, false>(s_cen, c, r, gz, gx, gy, okd, P, tab, sum);
                if (abort_test(sum, (double)(r + 1) * fs, A)) return 1e9f;
            }
        }
    }
    return sum / (25.0f * (float)nsrc);
}

// Correctly rounded a / b from the hardware reciprocal: one Newton step on
// 1/b, then two FMA quotient corrections -- the compiler's IEEE division
// sequence without its operand scaling (v_div_scale) and special-case fixup
// (v_div_fixup).  Both are identities when b and 1/b are normal, a / b is
// normal, |a| >= 2^-104 and exponent(a) - exponent(b) < 96, so the result
// then equals a / b bit for bit.  For |a| < 2^-104 the quotient can be off
// by a few units of a number below 2^-94, which cannot move
// trunc(f * x/z + c) and so no census index.  The host enables this only for
// jobs whose bounds keep every valid tap inside that domain (JobDesc.pinhole_ok
// bit 1: zmin >= 1e-6, zmax < 1e6, |ray| < 1e3, |ts| < 1e12).
__device__ __forceinline__ float div_rn_fast(float a, float b) {
    float r = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, r, 1.0f);
    r = __builtin_fmaf(e, r, r);
    float q = a * r;
    float t = __builtin_fmaf(-b, q, a);
    q = __builtin_fmaf(t, r, q);
    t = __builtin_fmaf(-b, q, a);
    return __builtin_fmaf(t, r, q);
}

// Fourth evaluator (same values): no hoisted window geometry, per-tap
// geometry recomputed for every source (an evaluation rarely gets past its
// first source, so holding 75 hoisted registers would only cost occupancy).
// Divides use div_rn_fast; fronto-parallel planes (the init sweep) take
// z = d, exact because (0 * rx + 0 * ry) + 1 is 1 for finite rays.
//
// Strengthened abort, same values: every tap adds a non-negative amount, so
// the f32 running sum never decreases and RN(sum / D) is monotone in sum.
// The first row of source s+1 tests sum / (5 (s+2)), a smaller divisor than
// the later rows of source s use (e.g. 10 against 15, 20, 25 for s = 0).  So
// once RN(sum / 5 (s+2)) > thr the evaluation is certain to abort and its
// value is 1e9 whichever test fires first: each row end tests against the
// smallest divisor of any test still ahead.  With PAIR, rows 1 and 2 of
// source 0 issue their ten gathers together.
template <bool FRONTO, bool PIN, bool FD = true>
__device__ __forceinline__ bool tap_v4(float2 ray, float nx, float ny, float nz, float d, const SrcP& P, unsigned& off) {
    float z, xr, yr;
    bool vd;
    if (FRONTO) {
        z = d;
        vd = true;
    } else {
        const float denom = (nx * ray.x + ny * ray.y) + nz;
        vd = !(denom < 0.05f);
        z = FD ? div_rn_fast(d, denom) : d / denom;
    }
    xr = z * ray.x;
    yr = z * ray.y;
    const float xs = ((P.r0 * xr + P.r1 * yr) + P.r2 * z) + P.t0;
    const float ys = ((P.r3 * xr + P.r4 * yr) + P.r5 * z) + P.t1;
    const float zs = ((P.r6 * xr + P.r7 * yr) + P.r8 * z) + P.t2;
    const float xn = FD ? div_rn_fast(xs, zs) : xs / zs, yn = FD ? div_rn_fast(ys, zs) : ys / zs;
    float fu, fv;
    if (PIN) {
        fu = P.f * xn + P.cx;
        fv = P.f * yn + P.cy;
    } else {
        const float dist = 1.0f + P.k * (xn * xn + yn * yn);
        fu = P.f * xn * dist + P.cx;
        fv = P.f * yn * dist + P.cy;
    }
    const int u = cvt_i32_rust(fu), v = cvt_i32_rust(fv);
    const bool ok = vd && !(zs < 1e-3f) && ((unsigned)u < (unsigned)P.sw) && ((unsigned)v < (unsigned)P.sh);
    // u < sw < 2^24 and v < sh < 2^24 (host check), so the 24-bit multiply is exact
    off = ok ? (__umul24((unsigned)v, (unsigned)P.sw) + (unsigned)u) * 8u : 0u;   // byte offset of the census word
    return ok;
}

template <int PW, bool FRONTO, bool PIN>
__device__ __forceinline__ void row_v4(const u64* __restrict__ s_cen, const float2* __restrict__ s_ray, int c, int r,
                                       float nx, float ny, float nz, float d, const SrcP& P,
                                       const float* __restrict__ tab, float& sum) {
    unsigned off[5];
    bool ok[5];
#pragma unroll
    for (int q = 0; q < 5; ++q)
        ok[q] = tap_v4<FRONTO, PIN>(s_ray[c + (r - 2) * PW + (q - 2)], nx, ny, nz, d, P, off[q]);
    u64 cs[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) cs[q] = cen_at(P, off[q]);
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        const u64 rc = s_cen[c + (r - 2) * PW + (q - 2)];
        // tab[48] = 48 / 48 = 1 exactly: an invalid tap's 1.0 from the same
        // table read; the popcount is pinned (empty asm) so that it is
        // computed for every lane and selected, with no branch around it
        unsigned pc = (unsigned)__popcll(rc ^ cs[q]);
        asm volatile("" : "+v"(pc));
        sum += tab[ok[q] ? pc : 48u];
    }
}

// Source-0 rows 1 and 2 with their ten gathers issued together; the two
// row tests are applied in order afterwards (1e9 iff either fires).
template <int PW, bool FRONTO, bool PIN>
__device__ __forceinline__ bool rows12_v4(const u64* __restrict__ s_cen, const float2* __restrict__ s_ray, int c,
                                          float nx, float ny, float nz, float d, const SrcP& P,
                                          const float* __restrict__ tab, const AbortThr& A, double D1, double D2,
                                          float& sum) {
    unsigned off[10];
    unsigned okm = 0;
#pragma unroll
    for (int t = 0; t < 10; ++t) {
        const int r = 1 + t / 5, q = t % 5;
        const bool ok = tap_v4<FRONTO, PIN>(s_ray[c + (r - 2) * PW + (q - 2)], nx, ny, nz, d, P, off[t]);
        okm |= (ok ? 1u : 0u) << t;
    }
    u64 cs[10];
#pragma unroll
    for (int t = 0; t < 10; ++t) cs[t] = cen_at(P, off[t]);
    bool ab = false;
#pragma unroll
    for (int t = 0; t < 10; ++t) {
        const int r = 1 + t / 5, q = t % 5;
        const u64 rc = s_cen[c + (r - 2) * PW + (q - 2)];
        unsigned pc = (unsigned)__popcll(rc ^ cs[t]);
        asm volatile("" : "+v"(pc));
        sum += tab[((okm >> t) & 1u) ? pc : 48u];
        if (t == 4) ab = abort_test(sum, D1, A);
        if (t == 9) ab = ab || abort_test(sum, D2, A);
    }
    return ab;
}

template <int PW, bool FRONTO, bool PIN, bool PAIR>
__device__ __forceinline__ float eval_v4_body(const u64* __restrict__ s_cen, const float2* __restrict__ s_ray, int c,
                                              const SrcDesc* __restrict__ S, int nsrc, float nx, float ny, float nz,
                                              float d, const AbortThr& A, const float* __restrict__ tab, int& nr) {
    float sum = 0.0f;
    {
        // divisors of source 0's row tests 5 (r + 1), capped at 10 by source
        // 1's first test; selected as constants (wave-uniform, scalar) rather
        // than formed with a vector min, whose results were spilled
        const bool capped = nsrc > 1;
        const SrcP P = load_src(S[0]);
        row_v4<PW, FRONTO, PIN>(s_cen, s_ray, c, 0, nx, ny, nz, d, P, tab, sum);
        nr = 1;
        if (abort_test(sum, 5.0, A)) return 1e9f;
        if (PAIR) {
            nr = 3;
            if (rows12_v4<PW, FRONTO, PIN>(s_cen, s_ray, c, nx, ny, nz, d, P, tab, A, 10.0, capped ? 10.0 : 15.0,
                                           sum))
                return 1e9f;
        } else {
#pragma unroll
            for (int r = 1; r < 3; ++r) {
                row_v4<PW, FRONTO, PIN>(s_cen, s_ray, c, r, nx, ny, nz, d, P, tab, sum);
                nr = r + 1;
                if (abort_test(sum, capped ? 10.0 : (double)(5 * (r + 1)), A)) return 1e9f;
            }
        }
#pragma unroll
        for (int r = 3; r < 5; ++r) {
            row_v4<PW, FRONTO, PIN>(s_cen, s_ray, c, r, nx, ny, nz, d, P, tab, sum);
            nr = r + 1;
            if (abort_test(sum, capped ? 10.0 : (double)(5 * (r + 1)), A)) return 1e9f;
        }
    }
    // not unrolled: evaluations rarely reach a second source, and unrolling by
    // the run-time source count multiplies the kernel's code about tenfold
#pragma nounroll
    for (int si = 1; si < nsrc; ++si) {
        // si is the same in every lane still here; saying so keeps the
        // source parameters in scalar registers (a loop with lane-dependent
        // exits is otherwise treated as divergent and loads them per lane)
        const int su = __builtin_amdgcn_readfirstlane(si);
        const SrcP P = load_src(S[su]);
        // row r's divisor 5 (r + 1) (si + 1), capped at the next source's
        // first divisor 5 (si + 2) (exact integers, as fmin of the products)
        const bool capped = su + 1 < nsrc;
        // an opaque copy of the window index keeps the compiler from hoisting
        // the 25 ray and 25 census LDS reads out of this loop (100 registers)
        int cc = c;
        asm volatile("" : "+v"(cc));
        if (P.k == 0.0f) {
#pragma unroll
            for (int r = 0; r < 5; ++r) {
                row_v4<PW, FRONTO, true>(s_cen, s_ray, cc, r, nx, ny, nz, d, P, tab, sum);
                nr = 5 * si + r + 1;
                const int dv = (r == 0 || !capped) ? (r + 1) * (su + 1) : su + 2;
                if (abort_test(sum, (double)(5 * dv), A)) return 1e9f;
            }
        } else {
#pragma unroll
            for (int r = 0; r < 5; ++r) {
                row_v4<PW, FRONTO, false>(s_cen, s_ray, cc, r, nx, ny, nz, d, P, tab, sum);
                nr = 5 * si + r + 1;
                const int dv = (r == 0 || !capped) ? (r + 1) * (su + 1) : su + 2;
                if (abort_test(sum, (double)(5 * dv), A)) return 1e9f;
            }
        }
    }
    nr = 5 * nsrc;
    return sum / (25.0f * (float)nsrc);
}

// Entry of the fourth evaluator; jobs outside the fast-divide domain (bit 1
// of pinhole_ok clear) take the reference-order evaluator, which holds no
// hoisted state either.  `nr` receives the number of window rows the
// evaluation read (5 per source in full; the reference-order fallback
// reports its full count), a measure of its cost used for scheduling only.
template <int PW, bool FRONTO, bool PAIR>
__device__ __forceinline__ float eval_plane_v4(const u64* __restrict__ s_cen, const float2* __restrict__ s_ray, int c,
                                               const SrcDesc* __restrict__ S, int nsrc, float nx, float ny, float nz,
                                               float d, float thr, const float* __restrict__ tab, int pin_flags,
                                               int& nr) {
    if (!(pin_flags & 2)) {
        nr = 5 * nsrc;
        return eval_plane<false>(LdsRef{s_cen, s_ray, c, PW}, S, nsrc, nx, ny, nz, d, thr, nullptr);
    }
    const AbortThr A = make_abort_thr(thr);
    if (S[0].k == 0.0f)
        return eval_v4_body<PW, FRONTO, true, PAIR>(s_cen, s_ray, c, S, nsrc, nx, ny, nz, d, A, tab, nr);
    return eval_v4_body<PW, FRONTO, false, PAIR>(s_cen, s_ray, c, S, nsrc, nx, ny, nz, d, A, tab, nr);
}

__device__ __forceinline__ void fill_pc_table(float* tab) {
    if (threadIdx.x < 64) tab[threadIdx.x] = pc_over_48((int)threadIdx.x);
}

// --------------------------------------------------------------------------
// tile staging
// --------------------------------------------------------------------------

// Global-address-space loads (not flat ones, which also count against the
// LDS wait counter).
template <int TW, int TH, int NT = TW * TH>
__device__ __forceinline__ void stage_tile(const JobDesc& J, int x0, int y0, u64* s_cen, float2* s_ray) {
    constexpr int PW = TW + 4, PH = TH + 4;
    gu64* cen = (gu64*)J.census;
    gu64* ray = (gu64*)J.rays;   // a float2 moved as its 8 bytes
    u64* s_ray8 = reinterpret_cast<u64*>(s_ray);
    const int w = J.w, h = J.h;
    for (int i = threadIdx.x; i < PW * PH; i += NT) {
        const int ay = i / PW, ax = i % PW;
        const int gx = min(max(x0 - 2 + ax, 0), w - 1);
        const int gy = min(max(y0 - 2 + ay, 0), h - 1);
        const size_t g = (size_t)gy * w + gx;
        s_cen[i] = cen[g];
        s_ray8[i] = ray[g];
    }
}

// XCD-aware block order: hardware hands consecutive workgroups to the 8
// XCDs round-robin; remapping gives each XCD a contiguous run of tiles so
// neighbouring tiles (which gather overlapping source-census lines) share
// that XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int n) {
    const int x = b & 7, i = b >> 3;
    const int per = n >> 3, rem = n & 7;
    return x * per + min(x, rem) + i;
}

// --------------------------------------------------------------------------
// init sweep (stereo.rs:247-269)
// --------------------------------------------------------------------------

template <int TW, int TH, int EV>
__device__ __forceinline__ void init_tile(const JobDesc& J, const SrcDesc* __restrict__ srcs, int x0, int y0, int nh,
                                          u64* s_cen, float2* s_ray, unsigned& ne, const float* s_tab) {
    constexpr int PW = TW + 4;
    const int lx = threadIdx.x % TW, ly = threadIdx.x / TW;
    const int x = x0 + lx, y = y0 + ly;
    if (x >= J.w || y >= J.h) return;
    const LdsRef R{s_cen, s_ray, (ly + 2) * PW + (lx + 2), PW};
    const SrcDesc* S = srcs + J.src_off;
    const float inv_min = J.inv_min, inv_max = J.inv_max;
    float best_c = __builtin_inff(), best_d = 0.0f;
    for (int hi = 0; hi < nh; ++hi) {
        const float inv = inv_min + (inv_max - inv_min) * (float)hi / (float)(nh - 1);
        const float d = 1.0f / inv;
        const float thr = (best_c + 0.02f) + 0.12f;
        int nr;
        const float c = EV >= 3 ? eval_plane_v4<PW, true, EV == 3>(s_cen, s_ray, R.c, S, J.nsrc, 0.0f, 0.0f, 1.0f, d,
                                                                  thr, s_tab, J.pinhole_ok, nr)
                        : EV == 2 ? eval_plane_v3<PW>(s_cen, s_ray, R.c, S, J.nsrc, 0.0f, 0.0f, 1.0f, d, thr, s_tab,
                                                      J.pinhole_ok != 0)
                        : EV == 1 ? eval_plane_fast<PW>(s_cen, s_ray, R.c, S, J.nsrc, 0.0f, 0.0f, 1.0f, d, thr)
                                  : eval_plane<false>(R, S, J.nsrc, 0.0f, 0.0f, 1.0f, d, thr, nullptr);
        ++ne;
        if (c < best_c) { best_c = c; best_d = d; }
    }
    const size_t i = (size_t)y * J.w + x;
    J.plane[0][i] = make_float4(0.0f, 0.0f, 1.0f, best_d);
    J.cost[0][i] = fminf(best_c, 1.0f);
}

// Init sweep with shared tap costs (evaluators 3 and 4), same values as
// init_tile.  The 48 hypotheses are fronto-parallel, so a window tap lies at
// depth d whichever window it belongs to and its term of plane_cost --
// tab[popcount] at a valid tap, 1 = tab[48] at an invalid one -- depends on
// the tap's own pixel only.  Per hypothesis and source the block computes
// the code of every pixel of its tile and apron once into LDS; each pixel
// then adds its 25 terms in plane_cost's order (source, row, column) with
// the same (strengthened) abort tests; LDS holds each tap's term (the /48
// table entry of its popcount, 1 when invalid), looked up once when the
// code is made instead of at each of the up to 25 windows that read it.
// Every hypothesis needs source 0 at
// every pixel, so the source-0 codes of INIT_NB hypotheses are computed
// together (their census gathers in flight at once, one barrier for all);
// a later source is costed one hypothesis at a time, while some pixel of
// the block still needs it (block vote).  `want` selects the pixels that
// run the sweep (all of them, or those the learned initialisation did not
// settle); every thread of the block must call this.
constexpr int INIT_NB = 2;

template <bool PIN, bool FD>
__device__ __forceinline__ unsigned fronto_code(u64 cref, float2 ray, float d, const SrcP& P) {
    unsigned off;
    const bool ok = tap_v4<true, PIN, FD>(ray, 0.0f, 0.0f, 1.0f, d, P, off);
    return ok ? (unsigned)__popcll(cref ^ cen_at(P, off)) : 48u;
}

// One source's terms of the tile + apron, this thread's pixels unrolled so
// that their census gathers are in flight together.
template <int NP, int NT, bool PIN, bool FD>
__device__ __forceinline__ void box_codes(const u64* s_cen, const float2* s_ray, float d, const SrcP& P,
                                          const float* tab, float* s_term) {
#pragma unroll
    for (int k = 0; k < (NP + NT - 1) / NT; ++k) {
        const int i = (int)threadIdx.x + k * NT;
        if (i < NP) s_term[i] = tab[fronto_code<PIN, FD>(s_cen[i], s_ray[i], d, P)];
    }
}

// Source-0 terms of hypotheses d[0 .. nb) into s_term[b * NP ..]: every
// projection first, then every gather, then the popcounts.  The nb * NP
// (hypothesis, tap) items are dealt to the threads as one list, so an 8x8
// tile's 2 x 144 items take 5 rounds of 64 rather than 2 x 3.
template <int NP, int NT, bool PIN, bool FD>
__device__ __forceinline__ void box_codes_batch(const u64* s_cen, const float2* s_ray, const float* d, int nb,
                                                const SrcP& P, const float* tab, float* s_term) {
    constexpr int K = (INIT_NB * NP + NT - 1) / NT;
    const int n = nb * NP;
    unsigned off[K];
    bool ok[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int j = (int)threadIdx.x + k * NT;
        const int b = j / NP, i = j - b * NP;
        float db = d[0];   // d[b] as selects (no indexed register array)
#pragma unroll
        for (int t = 1; t < INIT_NB; ++t)
            if (b == t) db = d[t];
        ok[k] = false;
        off[k] = 0;
        if (j < n) ok[k] = tap_v4<true, PIN, FD>(s_ray[i], 0.0f, 0.0f, 1.0f, db, P, off[k]);
    }
    u64 cs[K];
#pragma unroll
    for (int k = 0; k < K; ++k) cs[k] = ok[k] ? cen_at(P, off[k]) : 0ull;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int j = (int)threadIdx.x + k * NT;
        const int i = j - (j / NP) * NP;
        if (j < n) s_term[j] = tab[ok[k] ? (unsigned)__popcll(s_cen[i] ^ cs[k]) : 48u];
    }
}

// s_term: room for (INIT_NB + 1) * (TW + 4) * (TH + 4) terms.
template <int TW, int TH>
__device__ __forceinline__ void init_box_tile(const JobDesc& J, const SrcDesc* __restrict__ srcs, int x0, int y0,
                                              int nh, bool want, const u64* s_cen, const float2* s_ray,
                                              float* s_term, unsigned& ne, const float* s_tab) {
    constexpr int PW = TW + 4, NP = (TW + 4) * (TH + 4), NT = TW * TH;
    const int lx = threadIdx.x % TW, ly = threadIdx.x / TW;
    const int x = x0 + lx, y = y0 + ly;
    const bool mine = want && x < J.w && y < J.h;
    if (!__syncthreads_or(mine)) return;
    const int c = (ly + 2) * PW + (lx + 2);
    const SrcDesc* S = srcs + J.src_off;
    const int nsrc = J.nsrc;
    const bool fast = (J.pinhole_ok & 2) != 0, pin_ok = (J.pinhole_ok & 1) != 0;
    const float inv_min = J.inv_min, inv_max = J.inv_max;
    const SrcP P0 = load_src(S[0]);
    const bool pin0 = pin_ok && P0.k == 0.0f;
    float* s_later = s_term + INIT_NB * NP;   // terms of sources >= 1
    float best_c = __builtin_inff(), best_d = 0.0f;
    for (int h0 = 0; h0 < nh; h0 += INIT_NB) {
        const int nb = min(INIT_NB, nh - h0);
        float d[INIT_NB];
#pragma unroll
        for (int b = 0; b < INIT_NB; ++b) {
            const int hi = min(h0 + b, nh - 1);
            const float inv = inv_min + (inv_max - inv_min) * (float)hi / (float)(nh - 1);
            d[b] = 1.0f / inv;
        }
        __syncthreads();   // the previous batch's sums are done with s_term
        if (fast && pin0)
            box_codes_batch<NP, NT, true, true>(s_cen, s_ray, d, nb, P0, s_tab, s_term);
        else if (fast)
            box_codes_batch<NP, NT, false, true>(s_cen, s_ray, d, nb, P0, s_tab, s_term);
        else if (pin0)
            box_codes_batch<NP, NT, true, false>(s_cen, s_ray, d, nb, P0, s_tab, s_term);
        else
            box_codes_batch<NP, NT, false, false>(s_cen, s_ray, d, nb, P0, s_tab, s_term);
        __syncthreads();
        for (int b = 0; b < nb; ++b) {
            const float dd = d[b];
            const float thr = (best_c + 0.02f) + 0.12f;
            const AbortThr A = make_abort_thr(thr);
            float sum = 0.0f;
            bool live = mine, aborted = false;
            // the 25 terms of source si from `terms`, row tests as in
            // eval_v4_body: row r's divisor 5 (r + 1) (si + 1), capped at the
            // next source's first divisor 5 (si + 2), selected as integers
            auto add_source = [&](const float* terms, int si) {
                const bool capped = si + 1 < nsrc;
#pragma unroll
                for (int r = 0; r < 5; ++r) {
#pragma unroll
                    for (int q = 0; q < 5; ++q) sum += terms[c + (r - 2) * PW + (q - 2)];
                    const int dv = (r == 0 || !capped) ? (r + 1) * (si + 1) : si + 2;
                    if (abort_test(sum, (double)(5 * dv), A)) {
                        aborted = true;
                        live = false;
                        break;
                    }
                }
                if (si + 1 == nsrc) live = false;
            };
            if (live) add_source(s_term + b * NP, 0);
#pragma nounroll
            for (int si = 1; si < nsrc; ++si) {
                // the vote is also the barrier after the previous source's sums
                if (!__syncthreads_or(live)) break;
                // si is the same in every lane (the vote is block-wide); saying
                // so keeps the source parameters and census descriptor scalar
                const SrcP P = load_src(S[__builtin_amdgcn_readfirstlane(si)]);
                const bool pin = pin_ok && P.k == 0.0f;
                if (fast && pin)
                    box_codes<NP, NT, true, true>(s_cen, s_ray, dd, P, s_tab, s_later);
                else if (fast)
                    box_codes<NP, NT, false, true>(s_cen, s_ray, dd, P, s_tab, s_later);
                else if (pin)
                    box_codes<NP, NT, true, false>(s_cen, s_ray, dd, P, s_tab, s_later);
                else
                    box_codes<NP, NT, false, false>(s_cen, s_ray, dd, P, s_tab, s_later);
                __syncthreads();
                if (live) add_source(s_later, si);
            }
            if (mine) {
                ++ne;
                const float cst = aborted ? 1e9f : sum / (25.0f * (float)nsrc);
                if (cst < best_c) {
                    best_c = cst;
                    best_d = dd;
                }
            }
        }
    }
    if (mine) {
        const size_t i = (size_t)y * J.w + x;
        J.plane[0][i] = make_float4(0.0f, 0.0f, 1.0f, best_d);
        J.cost[0][i] = fminf(best_c, 1.0f);
    }
}

// --------------------------------------------------------------------------
// one pass (stereo.rs:279-369)
// --------------------------------------------------------------------------

// Neighbour candidate q = 0..3 (L, R, U, D) of pixel (x, y), stereo.rs:
// 293-311: the neighbour's plane carried to this pixel.  ray_at(dx, dy)
// gives the ray of the pixel at offset (dx, dy).  Returns false where the
// reference skips it.
template <class RayAt>
__device__ __forceinline__ bool nb_candidate(int q, int x, int y, int w, int h, const gplane* prev, RayAt ray_at,
                                             float2 ri, float zlo, float zhi, float& cnx, float& cny, float& cnz,
                                             float& d_plane) {
    const int dx = q == 0 ? -1 : (q == 1 ? 1 : 0);
    const int dy = q == 2 ? -1 : (q == 3 ? 1 : 0);
    const int xx = x + dx, yy = y + dy;
    if (xx < 0 || yy < 0 || xx >= w || yy >= h) return false;
    const float4 nb = as_float4(prev[(size_t)yy * w + xx]);
    const float2 rj = ray_at(dx, dy);
    const float a_nb = (nb.x * rj.x + nb.y * rj.y) + nb.z;
    if (a_nb < 0.1f) return false;
    d_plane = nb.w * a_nb;
    const float a_here = (nb.x * ri.x + nb.y * ri.y) + nb.z;
    if (a_here < 0.1f) return false;
    const float zc = d_plane / a_here;
    if (!(zlo <= zc && zc <= zhi)) return false;
    cnx = nb.x;
    cny = nb.y;
    cnz = nb.z;
    return true;
}

// The neighbour candidates of a pixel, 2 bits each (candidate q at bits
// 2q): 0 = skipped by the reference, 1 = a repeat (counted, not
// evaluated), 2 = evaluated.
//
// A repeat is a candidate whose plane's cost here is already accounted
// for; it cannot change the result.  plane_cost is a function of
// (n, d_plane) and the hint, returning the plane's cost C or 1e9 when the
// hint aborts it, and the best cost only falls: a repeat's threshold is no
// higher than the earlier evaluation's, so its result is never below that
// one's, and best is already at most that one's result.  So the repeat is
// never accepted.  The planes compared against are
//   * the current plane (n, cur_d), cur_d the d_plane it was costed with:
//     every plane is stored together with min(C, 1) for its cost C here at
//     that d_plane (by the init sweep, an accepted candidate or the learned
//     path's proposal costs; stereo.rs:263-266, 355-366), so best <=
//     cur_cost <= C <= 1 < 1e9 from the start.  Only with costs_exact:
//     planted states carry arbitrary costs;
//   * the first neighbour evaluated here (best <= its result after it).
// No draws are involved.  Neighbours mostly repeat the current plane: the
// init's fronto-parallel planes persist wherever no candidate wins, and a
// fronto-parallel neighbour carries d_plane = z * 1 exactly.
template <class RayAt>
__device__ __forceinline__ unsigned nb_codes(int x, int y, int w, int h, const gplane* prev, RayAt ray_at, float2 ri,
                                             float4 cur, float cur_d, float zlo, float zhi, bool costs_exact) {
    const uint32_t c0 = __float_as_uint(cur.x), c1 = __float_as_uint(cur.y), c2 = __float_as_uint(cur.z),
                   c3 = __float_as_uint(cur_d);
    uint32_t f0 = 0u, f1 = 0u, f2 = 0u, f3 = 0u;
    bool have_first = false;
    unsigned codes = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float cnx, cny, cnz, d_plane;
        if (!nb_candidate(q, x, y, w, h, prev, ray_at, ri, zlo, zhi, cnx, cny, cnz, d_plane)) continue;
        const uint32_t u0 = __float_as_uint(cnx), u1 = __float_as_uint(cny), u2 = __float_as_uint(cnz),
                       u3 = __float_as_uint(d_plane);
        if ((costs_exact && u0 == c0 && u1 == c1 && u2 == c2 && u3 == c3) ||
            (have_first && u0 == f0 && u1 == f1 && u2 == f2 && u3 == f3)) {
            codes |= 1u << (2 * q);
            continue;
        }
        codes |= 2u << (2 * q);
        if (!have_first) {
            f0 = u0;
            f1 = u1;
            f2 = u2;
            f3 = u3;
            have_first = true;
        }
    }
    return codes;
}

// Candidate q of pixel (x, y) in the reference's order: q = 0..3 the
// neighbour planes L, R, U, D carried to this pixel (nb_candidate), q = 4..6
// the perturbations of the current plane (stereo.rs:324-353), which draw
// from `seed`.  Returns false where the reference skips the candidate; the
// draws made stay consumed (a failed inv_new / zcand test consumes only g0).
template <class RayAt>
__device__ __forceinline__ bool make_candidate(int q, int x, int y, int w, int h, const gplane* prev, RayAt ray_at,
                                               float2 ri, float4 cur, u64& seed, float sig_inv, float sig_n,
                                               float range, float zlo, float zhi, float& cnx, float& cny, float& cnz,
                                               float& d_plane) {
    if (q < 4) return nb_candidate(q, x, y, w, h, prev, ray_at, ri, zlo, zhi, cnx, cny, cnz, d_plane);
    const float inv = 1.0f / fmaxf(cur.w, 1e-6f);
    const float g0 = gauss3(seed);
    const float inv_new = inv + sig_inv * range * g0;
    if (inv_new <= 0.0f) return false;
    const float zc = 1.0f / inv_new;
    if (!(zlo <= zc && zc <= zhi)) return false;
    const float g1 = gauss3(seed);
    float nnx = cur.x + sig_n * g1;
    const float g2 = gauss3(seed);
    float nny = cur.y + sig_n * g2;
    const float g3 = gauss3(seed);
    float nnz = cur.z + sig_n * g3;
    const float len = sqrtf((nnx * nnx + nny * nny) + nnz * nnz);
    if (len < 1e-6f) return false;
    nnx = nnx / len;
    nny = nny / len;
    nnz = nnz / len;
    const float a_here = (nnx * ri.x + nny * ri.y) + nnz;
    if (a_here < 0.1f) return false;
    d_plane = zc * a_here;
    cnx = nnx;
    cny = nny;
    cnz = nnz;
    return true;
}

// One pixel of one pass: pixel (x, y) at window centre c of the LDS tile of
// pitch PW.  With WORK, the window rows its evaluations read, halved and
// capped at 255, go to J.work (scheduling information only).
template <int PW, int EV, bool WORK = false>
__device__ __forceinline__ void pass_pixel(const JobDesc& J, const SrcDesc* __restrict__ S, int x, int y, int c,
                                          int par, u64 salt, float sig_inv, float sig_n, u64* s_cen, float2* s_ray,
                                          unsigned& ne, const float* s_tab) {
    const int w = J.w, h = J.h;
    const LdsRef R{s_cen, s_ray, c, PW};
    const int nsrc = J.nsrc;
    const size_t i = (size_t)y * w + x;
    // planes and costs of the previous pass through global-address-space
    // loads (flat ones would also hold up the LDS waits of the window reads)
    gplane* prev = (gplane*)J.plane[par];
    const float2 ri = s_ray[c];
    u64 seed = pixel_rng(x, y, salt);
    const float4 cur = as_float4(prev[i]);
    const float cur_cost = ((gfloat*)J.cost[par])[i];
    float best_c = cur_cost;
    float b0 = cur.x, b1 = cur.y, b2 = cur.z;
    float bd = cur.w * ((b0 * ri.x + b1 * ri.y) + b2);
    const float zlo = J.zmin * 0.5f, zhi = J.zmax * 2.0f;
    int rows = 0;
    auto ray_at = [&](int dx, int dy) { return s_ray[c + dy * PW + dx]; };
    // the d_plane the current plane was costed with: z for (+0, +0, 1, z),
    // else stored beside it (JobDesc.dplane)
    const bool exact = J.costs_exact != 0;
    const bool fronto = __float_as_uint(cur.x) == 0u && __float_as_uint(cur.y) == 0u &&
                        __float_as_uint(cur.z) == 0x3f800000u;
    const float cur_d = fronto || !exact ? cur.w : J.dplane[i];
    const unsigned codes = nb_codes(x, y, w, h, prev, ray_at, ri, cur, cur_d, zlo, zhi, exact);

    // Seven candidates in the reference's order: the 4 neighbour planes
    // L, R, U, D (stereo.rs:293-321), then 3 perturbations of the current
    // plane (stereo.rs:324-353).  Candidate geometry never depends on
    // costs, so one evaluation site serves all seven; each candidate's
    // hint is the best cost found before it.  Repeated neighbour planes are
    // counted without an evaluation (nb_codes), and each lane walks only its
    // pending candidates (the neighbours it evaluates, then the three
    // perturbations) in order: a round's lanes stay in lock step over the
    // i-th pending candidate of each, so a neighbour evaluation of one pixel
    // costs the round one step, not one per neighbour direction.
    ne += (unsigned)__popc(codes & 0x55u);   // repeats (code 1)
    unsigned pend = 0x70u;                   // q = 4, 5, 6
#pragma unroll
    for (int q = 0; q < 4; ++q) pend |= ((codes >> (2 * q + 1)) & 1u) << q;   // code 2
    const float range = J.inv_max - J.inv_min;
    while (pend) {
        const int q = __builtin_ctz(pend);
        pend &= pend - 1u;
        float cnx, cny, cnz, d_plane;
        if (!make_candidate(q, x, y, w, h, prev, ray_at, ri, cur, seed, sig_inv, sig_n, range, zlo, zhi, cnx, cny,
                            cnz, d_plane))
            continue;
        const float thr = (best_c + 0.02f) + 0.12f;
        int nr = 5 * nsrc;
        const float cst = EV >= 3 ? eval_plane_v4<PW, false, EV == 3>(s_cen, s_ray, c, S, nsrc, cnx, cny, cnz, d_plane,
                                                                      thr, s_tab, J.pinhole_ok, nr)
                          : EV == 2 ? eval_plane_v3<PW>(s_cen, s_ray, c, S, nsrc, cnx, cny, cnz, d_plane, thr, s_tab,
                                                        J.pinhole_ok != 0)
                          : EV == 1 ? eval_plane_fast<PW>(s_cen, s_ray, c, S, nsrc, cnx, cny, cnz, d_plane, thr)
                                    : eval_plane<false>(R, S, nsrc, cnx, cny, cnz, d_plane, thr, nullptr);
        ++ne;
        rows += nr;
        if (cst < best_c) { best_c = cst; b0 = cnx; b1 = cny; b2 = cnz; bd = d_plane; }
    }

    // accept on strict improvement (stereo.rs:355-367); otherwise carry over
    float4 outp = cur;
    float outc = cur_cost;
    if (best_c < cur_cost) {
        const float a_here = (b0 * ri.x + b1 * ri.y) + b2;
        if (a_here > 0.05f) {
            outp = make_float4(b0, b1, b2, bd / a_here);
            outc = fminf(best_c, 1.0f);
            J.dplane[i] = bd;   // this pixel's own entry: no other pixel reads it
        }
    }
    J.plane[par ^ 1][i] = outp;
    J.cost[par ^ 1][i] = outc;
    if (WORK && J.work) J.work[i] = (uint8_t)min((rows + 1) >> 1, 255);
}

template <int TW, int TH, int EV>
__device__ __forceinline__ void pass_tile(const JobDesc& J, const SrcDesc* __restrict__ srcs, int x0, int y0, int par,
                                          u64 salt, float sig_inv, float sig_n, u64* s_cen, float2* s_ray,
                                          unsigned& ne, const float* s_tab) {
    constexpr int PW = TW + 4;
    const int lx = threadIdx.x % TW, ly = threadIdx.x / TW;
    const int x = x0 + lx, y = y0 + ly;
    if (x >= J.w || y >= J.h) return;
    pass_pixel<PW, EV>(J, srcs + J.src_off, x, y, (ly + 2) * PW + (lx + 2), par, salt, sig_inv, sig_n, s_cen, s_ray,
                       ne, s_tab);
}

// --------------------------------------------------------------------------
// kernels: tile grid (one block per tile) and persistent tile queue
// --------------------------------------------------------------------------

template <int TW, int TH>
__device__ __forceinline__ void block_count(unsigned ne, unsigned* s_cnt, u64* counter) {
    atomicAdd(s_cnt, ne);
    __syncthreads();
    if (threadIdx.x == 0 && counter) atomicAdd(counter, (u64)*s_cnt);
}

template <int TW, int TH, int EV>
__global__ __launch_bounds__(TW* TH) void init_kernel(const JobDesc* __restrict__ jobs, const SrcDesc* __restrict__ srcs,
                                                      const int4* __restrict__ tiles, int ntiles, int nh,
                                                      u64* __restrict__ counter) {
    __shared__ u64 s_cen[(TW + 4) * (TH + 4)];
    __shared__ float2 s_ray[(TW + 4) * (TH + 4)];
    __shared__ float s_term[EV >= 3 ? (INIT_NB + 1) * (TW + 4) * (TH + 4) : 1];
    __shared__ unsigned s_cnt;
    __shared__ float s_tab[64];
    const int4 tl = tiles[xcd_remap(blockIdx.x, ntiles)];
    const JobDesc& J = jobs[tl.x];
    if (threadIdx.x == 0) s_cnt = 0;
    fill_pc_table(s_tab);
    stage_tile<TW, TH>(J, tl.y, tl.z, s_cen, s_ray);
    __syncthreads();
    unsigned ne = 0;
    if (EV >= 3)
        init_box_tile<TW, TH>(J, srcs, tl.y, tl.z, nh, true, s_cen, s_ray, s_term, ne, s_tab);
    else
        init_tile<TW, TH, EV>(J, srcs, tl.y, tl.z, nh, s_cen, s_ray, ne, s_tab);
    block_count<TW, TH>(ne, &s_cnt, counter);
}

// With the source parameters of sources >= 1 in scalar registers the
// per-tap evaluators (3, 4) need 89 registers: five waves per SIMD with no
// spill.  Held to 80 they run six waves with a 20-byte spill outside the
// tap loops, measured 1.8 % faster on the 8x8 shapes and 2-4 % on the
// 16x16 / 16x8 ones (same GPU, alternating runs) for 22 % more HBM traffic
// (spill slots; 9.5 -> 11.6 GB per pass launch, far below the HBM bound).
// Seven waves (72 registers, a 44-byte spill) measured no faster than six.
#define PASS_OCCUPANCY __attribute__((amdgpu_waves_per_eu(EV >= 3 ? 6 : 1, 8)))

template <int TW, int TH, int EV>
__global__ __launch_bounds__(TW* TH) PASS_OCCUPANCY void pass_kernel(const JobDesc* __restrict__ jobs, const SrcDesc* __restrict__ srcs,
                                                      const int4* __restrict__ tiles, int ntiles, int par, u64 salt,
                                                      float sig_inv, float sig_n, u64* __restrict__ counter) {
    __shared__ u64 s_cen[(TW + 4) * (TH + 4)];
    __shared__ float2 s_ray[(TW + 4) * (TH + 4)];
    __shared__ unsigned s_cnt;
    __shared__ float s_tab[64];
    const int4 tl = tiles[xcd_remap(blockIdx.x, ntiles)];
    const JobDesc& J = jobs[tl.x];
    if (threadIdx.x == 0) s_cnt = 0;
    fill_pc_table(s_tab);
    stage_tile<TW, TH>(J, tl.y, tl.z, s_cen, s_ray);
    __syncthreads();
    unsigned ne = 0;
    pass_tile<TW, TH, EV>(J, srcs, tl.y, tl.z, par, salt, sig_inv, sig_n, s_cen, s_ray, ne, s_tab);
    block_count<TW, TH>(ne, &s_cnt, counter);
}

// Persistent form: a grid sized to the machine pulls tiles from an atomic
// queue head (queue[0]); every block leaves once the head passes ntiles.
template <int TW, int TH, int EV>
__global__ __launch_bounds__(TW* TH) PASS_OCCUPANCY void pass_persistent_kernel(const JobDesc* __restrict__ jobs,
                                                                 const SrcDesc* __restrict__ srcs,
                                                                 const int4* __restrict__ tiles, int ntiles, int par,
                                                                 u64 salt, float sig_inv, float sig_n,
                                                                 u64* __restrict__ counter, int* __restrict__ queue) {
    __shared__ u64 s_cen[(TW + 4) * (TH + 4)];
    __shared__ float2 s_ray[(TW + 4) * (TH + 4)];
    __shared__ unsigned s_cnt;
    __shared__ int s_tile;
    __shared__ float s_tab[64];
    if (threadIdx.x == 0) s_cnt = 0;
    fill_pc_table(s_tab);
    unsigned ne = 0;
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(queue, 1);
        __syncthreads();
        // the same in every lane: saying so keeps the tile, job and source
        // descriptors in scalar registers
        const int t = __builtin_amdgcn_readfirstlane(s_tile);
        if (t >= ntiles) break;
        const int4 tl = tiles[t];
        const JobDesc& J = jobs[tl.x];
        stage_tile<TW, TH>(J, tl.y, tl.z, s_cen, s_ray);
        __syncthreads();
        pass_tile<TW, TH, EV>(J, srcs, tl.y, tl.z, par, salt, sig_inv, sig_n, s_cen, s_ray, ne, s_tab);
        __syncthreads();   // LDS tile and s_tile are reused by the next iteration
    }
    block_count<TW, TH>(ne, &s_cnt, counter);
}

// Sorted form: one wave per TW x TH tile (TW * TH / 64 rounds of 64 pixels).
// A pass's cost at a pixel is the window rows its seven evaluations read,
// and a wave's time on a candidate is that of its slowest lane; the rows a
// pixel read in the previous pass predict the current ones well (window
// texture and the plane's fit change slowly between passes).  So the tile's
// pixels are grouped by the previous pass's rows (J.work / 2, i.e. rows / 4,
// in 16 classes, the last open) and the rounds take them 64 at a time in
// class order, bringing pixels of similar cost into one wave.  A counting
// sort by wave ballots keeps each class in raster order.  Results do not
// depend on the order: every pixel reads only the previous pass's planes and
// the LDS window, and writes only its own outputs.
template <int TW, int TH, int EV>
__global__ __launch_bounds__(64) PASS_OCCUPANCY void pass_sorted_kernel(const JobDesc* __restrict__ jobs,
                                                                       const SrcDesc* __restrict__ srcs,
                                                                       const int4* __restrict__ tiles, int ntiles,
                                                                       int par, u64 salt, float sig_inv, float sig_n,
                                                                       u64* __restrict__ counter,
                                                                       int* __restrict__ queue) {
    constexpr int NPIX = TW * TH, NG = NPIX / 64, PW = TW + 4;
    constexpr unsigned NCLASS = 16;
    static_assert(NPIX % 64 == 0 && NPIX <= 1024, "a whole number of 64-pixel rounds");
    __shared__ u64 s_cen[(TW + 4) * (TH + 4)];
    __shared__ float2 s_ray[(TW + 4) * (TH + 4)];
    using perm_t = typename std::conditional<NPIX <= 256, uint8_t, uint16_t>::type;
    __shared__ perm_t s_perm[NPIX];
    __shared__ unsigned s_cnt;
    __shared__ int s_tile;
    __shared__ float s_tab[64];
    if (threadIdx.x == 0) s_cnt = 0;
    fill_pc_table(s_tab);
    unsigned ne = 0;
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(queue, 1);
        __syncthreads();
        // the same in every lane: saying so keeps the tile, job and source
        // descriptors in scalar registers
        const int t = __builtin_amdgcn_readfirstlane(s_tile);
        if (t >= ntiles) break;
        const int4 tl = tiles[t];
        const JobDesc& J = jobs[tl.x];
        const int w = J.w, h = J.h;
        stage_tile<TW, TH, 64>(J, tl.y, tl.z, s_cen, s_ray);
        // class of each of this lane's pixels (p = lane + 64 g); NCLASS = outside the view
        const uint8_t* wk = J.work;
        unsigned cls[NG];
        int n_in = 0;
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const int p = (int)threadIdx.x + 64 * g;
            const int x = tl.y + p % TW, y = tl.z + p / TW;
            cls[g] = NCLASS;
            if (x < w && y < h) cls[g] = wk ? min((unsigned)wk[(size_t)y * w + x] >> 1, NCLASS - 1) : 0u;
            n_in += __popcll(__ballot(cls[g] < NCLASS));
        }
        // counting sort: position = pixels of lower classes + those of the
        // same class earlier in (group, lane) order
        int pos[NG];
        int base = 0;
#pragma unroll
        for (unsigned v = 0; v < NCLASS; ++v) {
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                const u64 m = __ballot(cls[g] == v);
                const int below = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                pos[g] = cls[g] == v ? base + below : pos[g];
                base += __popcll(m);
            }
        }
#pragma unroll
        for (int g = 0; g < NG; ++g)
            if (cls[g] < NCLASS) s_perm[pos[g]] = (perm_t)(threadIdx.x + 64 * g);
        __syncthreads();
        const SrcDesc* S = srcs + J.src_off;
#pragma unroll 1
        for (int r = 0; r * 64 < n_in; ++r) {
            const int k = r * 64 + (int)threadIdx.x;
            if (k >= n_in) break;
            const int p = s_perm[k];
            const int lx = p % TW, ly = p / TW;
            pass_pixel<PW, EV, true>(J, S, tl.y + lx, tl.z + ly, (ly + 2) * PW + (lx + 2), par, salt, sig_inv, sig_n,
                                     s_cen, s_ray, ne, s_tab);
        }
        __syncthreads();   // LDS tile, order and s_tile are reused by the next iteration
    }
    block_count<TW, TH>(ne, &s_cnt, counter);
}

// Init sweep for the sorted form's tiles: one single-wave block per 8x8
// sub-tile, running init_box_tile as the 8x8 init launch does (same values
// as every other init launch).
template <int TW, int TH>
__global__ __launch_bounds__(64) void init_sub_kernel(const JobDesc* __restrict__ jobs,
                                                      const SrcDesc* __restrict__ srcs,
                                                      const int4* __restrict__ tiles, int ntiles, int nh,
                                                      u64* __restrict__ counter) {
    static_assert(TW % 8 == 0 && TH % 8 == 0, "whole 8x8 sub-tiles");
    constexpr int NSX = TW / 8, NSUB = NSX * (TH / 8);
    __shared__ u64 s_cen[12 * 12];
    __shared__ float2 s_ray[12 * 12];
    __shared__ float s_term[(INIT_NB + 1) * 12 * 12];
    __shared__ unsigned s_cnt;
    __shared__ float s_tab[64];
    const int b = xcd_remap(blockIdx.x, ntiles * NSUB);
    const int sub = b % NSUB;
    const int4 tl = tiles[b / NSUB];
    const JobDesc& J = jobs[tl.x];
    const int x0 = tl.y + (sub % NSX) * 8, y0 = tl.z + (sub / NSX) * 8;
    if (x0 >= J.w || y0 >= J.h) return;   // the same for the whole block
    if (threadIdx.x == 0) s_cnt = 0;
    fill_pc_table(s_tab);
    stage_tile<8, 8>(J, x0, y0, s_cen, s_ray);
    __syncthreads();
    unsigned ne = 0;
    init_box_tile<8, 8>(J, srcs, x0, y0, nh, true, s_cen, s_ray, s_term, ne, s_tab);
    block_count<8, 8>(ne, &s_cnt, counter);
}

// --------------------------------------------------------------------------
// gate (stereo.rs:373-393) + outputs
// --------------------------------------------------------------------------

template <int TW, int TH>
__global__ __launch_bounds__(TW* TH) void gate_kernel(const JobDesc* __restrict__ jobs, const int4* __restrict__ tiles,
                                                      int par, float cost_thr) {
    __shared__ unsigned s_valid;
    const int4 tl = tiles[blockIdx.x];
    const JobDesc& J = jobs[tl.x];
    if (threadIdx.x == 0) s_valid = 0;
    __syncthreads();
    const int x = tl.y + threadIdx.x % TW, y = tl.z + threadIdx.x / TW;
    if (x < J.w && y < J.h) {
        const size_t i = (size_t)y * J.w + x;
        const float4 p = J.plane[par][i];
        const float cst = J.cost[par][i];
        const int tex = __popcll(J.census[i]);
        const bool textured = tex >= 4 && tex <= 44;
        const bool ok = textured && cst < cost_thr && p.w > J.zmin * 0.5f && p.w < J.zmax * 2.0f;
        J.depth_out[i] = ok ? p.w : 0.0f;
        J.cost_out[i] = cst;
        if (J.normal_out) {
            J.normal_out[3 * i] = p.x;
            J.normal_out[3 * i + 1] = p.y;
            J.normal_out[3 * i + 2] = p.z;
        }
        // the reference counts depth > 0 (lib.rs:361)
        if (ok && p.w > 0.0f) atomicAdd(&s_valid, 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0 && J.valid_count && s_valid) atomicAdd(J.valid_count, (u64)s_valid);
}

// --------------------------------------------------------------------------
// learned initialisation support
// --------------------------------------------------------------------------

__device__ __forceinline__ uint16_t f2bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Fronto-parallel census sweep volume on a coarse grid: coarse pixel
// (cx, cy) is full-resolution pixel (S*cx, S*cy) with its full-resolution
// 5x5 window (clamped at the image border); nplanes planes uniform in inverse
// depth over [1/zmax, 1/zmin] (the init sweep's spacing with nplanes
// hypotheses).  The value of a plane is the window's census cost summed
// exactly in integers over the job's sources -- the Hamming distances of the
// valid taps plus 48 per invalid tap, i.e. plane_cost's Hamming sum and
// invalid count with no abort -- divided once by 48 * 25 * K.  On a
// fronto-parallel plane every tap lies at depth d whichever window it belongs
// to, so a pixel's tap cost is computed once per plane and shared by the 25
// windows that contain it: each plane costs the tile and its apron into LDS,
// then each coarse pixel sums its window there.  fp32 values (optional) and
// bf16 network features (value - 0.5) * 2 are written per coarse pixel.
template <bool PIN, bool FD>
__device__ __forceinline__ unsigned fronto_tap_cost(u64 cref, float2 ray, float d, const SrcP& P) {
    unsigned off;
    const bool ok = tap_v4<true, PIN, FD>(ray, 0.0f, 0.0f, 1.0f, d, P, off);
    return ok ? (unsigned)__popcll(cref ^ cen_at(P, off)) : 48u;
}

// One source's tap costs of this thread's pixels of the tile + apron.
template <int NP, int NT, bool PIN, bool FD>
__device__ __forceinline__ void vol_costs(const u64* s_cen, const float2* s_ray, float d, const SrcP& P,
                                          unsigned* tot) {
#pragma unroll
    for (int k = 0; k < (NP + NT - 1) / NT; ++k) {
        const int i = (int)threadIdx.x + k * NT;
        if (i < NP) tot[k] += fronto_tap_cost<PIN, FD>(s_cen[i], s_ray[i], d, P);
    }
}

// Four pinhole sources on the scaling-free divide (the usual case): the tap
// offsets of all four sources and the thread's NIT pixels first, then all
// their census gathers (4 NIT in flight at once), then the popcounts.
template <int NP, int NT>
__device__ __forceinline__ void vol_costs4(const u64* s_cen, const float2* s_ray, float d, const SrcP* P,
                                           unsigned* tot) {
    constexpr int K = (NP + NT - 1) / NT;
    unsigned off[4][K];
    bool ok[4][K];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int i = (int)threadIdx.x + k * NT;
            ok[s][k] = false;
            off[s][k] = 0;
            if (i < NP) ok[s][k] = tap_v4<true, true, true>(s_ray[i], 0.0f, 0.0f, 1.0f, d, P[s], off[s][k]);
        }
    u64 cs[4][K];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int k = 0; k < K; ++k) cs[s][k] = cen_at(P[s], off[s][k]);   // offset 0 when invalid: in bounds
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = (int)threadIdx.x + k * NT;
        if (i < NP) {
            const u64 cr = s_cen[i];
#pragma unroll
            for (int s = 0; s < 4; ++s) tot[k] += ok[s][k] ? (unsigned)__popcll(cr ^ cs[s][k]) : 48u;
        }
    }
}

template <int TW, int TH, int S>
__global__ __launch_bounds__(TW* TH) void volume_kernel(const JobDesc* __restrict__ jobs,
                                                        const SrcDesc* __restrict__ srcs,
                                                        const int4* __restrict__ tiles, int ntiles, int nplanes,
                                                        const VolOut* __restrict__ outs) {
    constexpr int PW = S * TW + 4, PH = S * TH + 4, NP = PW * PH, NT = TW * TH, NIT = (NP + NT - 1) / NT;
    __shared__ u64 s_cen[NP];
    __shared__ float2 s_ray[NP];
    __shared__ unsigned short s_cost[NP];
    const int4 tl = tiles[xcd_remap(blockIdx.x, ntiles)];
    const JobDesc& J = jobs[tl.x];
    const int w = J.w, h = J.h;
    const int fx0 = S * tl.y - 2, fy0 = S * tl.z - 2;
    for (int i = threadIdx.x; i < NP; i += TW * TH) {
        const int ay = i / PW, ax = i % PW;
        const int gx = min(max(fx0 + ax, 0), w - 1);
        const int gy = min(max(fy0 + ay, 0), h - 1);
        const size_t g = (size_t)gy * w + gx;
        s_cen[i] = J.census[g];
        s_ray[i] = J.rays[g];
    }
    const int cw = (w + S - 1) / S, ch = (h + S - 1) / S;
    const int lx = threadIdx.x % TW, ly = threadIdx.x / TW;
    const int x = tl.y + lx, y = tl.z + ly;
    const bool mine = x < cw && y < ch;
    const int c = (S * ly + 2) * PW + (S * lx + 2);
    const SrcDesc* Sd = srcs + J.src_off;
    const VolOut& O = outs[tl.x];
    const size_t oi = (size_t)y * cw + x;
    const bool fast = (J.pinhole_ok & 2) != 0, pin_ok = (J.pinhole_ok & 1) != 0;
    const float denom = (float)(48 * 25 * J.nsrc);
    bool four = J.nsrc == 4 && fast && pin_ok;
    for (int si = 0; si < 4 && four; ++si) four = Sd[si].k == 0.0f;
    for (int hi = 0; hi < nplanes; ++hi) {
        const float inv = J.inv_min + (J.inv_max - J.inv_min) * (float)hi / (float)(nplanes - 1);
        const float d = 1.0f / inv;
        __syncthreads();   // staging done / previous plane's sums read
        // sources outside, the thread's NIT pixels inside (unrolled): the
        // NIT census gathers of a source are in flight together; the
        // integer sums do not depend on the order
        unsigned tot[NIT];
#pragma unroll
        for (int k = 0; k < NIT; ++k) tot[k] = 0;
        if (four) {
            const SrcP P4[4] = {load_src(Sd[0]), load_src(Sd[1]), load_src(Sd[2]), load_src(Sd[3])};
            vol_costs4<NP, NT>(s_cen, s_ray, d, P4, tot);
        }
        for (int si = 0; si < (four ? 0 : J.nsrc); ++si) {
            const SrcP P = load_src(Sd[si]);
            const bool pin = pin_ok && P.k == 0.0f;
            if (fast && pin)
                vol_costs<NP, NT, true, true>(s_cen, s_ray, d, P, tot);
            else if (fast)
                vol_costs<NP, NT, false, true>(s_cen, s_ray, d, P, tot);
            else if (pin)
                vol_costs<NP, NT, true, false>(s_cen, s_ray, d, P, tot);
            else
                vol_costs<NP, NT, false, false>(s_cen, s_ray, d, P, tot);
        }
#pragma unroll
        for (int k = 0; k < NIT; ++k) {
            const int i = threadIdx.x + k * NT;
            if (i < NP) s_cost[i] = (unsigned short)tot[k];
        }
        __syncthreads();
        if (mine) {
            unsigned sum = 0;
#pragma unroll
            for (int dy = -2; dy <= 2; ++dy)
#pragma unroll
                for (int dx = -2; dx <= 2; ++dx) sum += s_cost[c + dy * PW + dx];
            const float v = (float)sum / denom;
            if (O.volf) O.volf[oi * nplanes + hi] = v;
            if (O.vol) O.vol[oi * 32 + hi] = f2bf16((v - 0.5f) * 2.0f);
        }
    }
}

// Initialisation from the proposal network: where the confidence reaches
// conf_thr and the proposed normal faces the ray (n . r >= 0.1), the pixel
// starts from the proposed plane with its plane_cost over all sources;
// elsewhere it runs the plain plane-sweep initialisation.
// One TW x TH tile (tl: job, x0, y0) of the learned initialisation; every
// thread of the block calls it.
template <int TW, int TH, int EV>
__device__ __forceinline__ void proposal_init_tile(const JobDesc* __restrict__ jobs, const SrcDesc* __restrict__ srcs,
                                                   const int4 tl, int nh, const float* const* __restrict__ preds,
                                                   int stride, float conf_thr, u64* __restrict__ counter,
                                                   u64* __restrict__ accepted, u64* __restrict__ list,
                                                   float4* __restrict__ prop) {
    constexpr int PW = TW + 4;
    __shared__ u64 s_cen[(TW + 4) * (TH + 4)];
    __shared__ float2 s_ray[(TW + 4) * (TH + 4)];
    __shared__ float s_term[EV >= 3 ? (INIT_NB + 1) * (TW + 4) * (TH + 4) : 1];
    __shared__ unsigned s_cnt, s_acc;
    __shared__ u64 s_base;
    __shared__ float s_tab[64];
    const JobDesc& J = jobs[tl.x];
    if (threadIdx.x == 0) {
        s_cnt = 0;
        s_acc = 0;
    }
    fill_pc_table(s_tab);
    stage_tile<TW, TH>(J, tl.y, tl.z, s_cen, s_ray);
    __syncthreads();
    unsigned ne = 0, na = 0;
    const int lx = threadIdx.x % TW, ly = threadIdx.x / TW;
    const int x = tl.y + lx, y = tl.z + ly;
    const bool inside = x < J.w && y < J.h;
    bool accept = false;
    float4 proposal = make_float4(0.f, 0.f, 0.f, 0.f);
    if (inside) {
        const int c = (ly + 2) * PW + (lx + 2);
        const size_t i = (size_t)y * J.w + x;
        const int cw = (J.w + stride - 1) / stride;
        const float* pr = preds[tl.x] + 5 * ((size_t)(y / stride) * cw + (x / stride));
        const float inv_n = pr[0], nx = pr[1], ny = pr[2], nz = pr[3], conf = pr[4];
        const float2 ri = s_ray[c];
        const float a = (nx * ri.x + ny * ri.y) + nz;
        // the proposed plane passes through depth 1/inv along the ray of the
        // coarse pixel (stride*(x/stride), stride*(y/stride)), which lies in
        // this tile because tile origins are even
        const float2 rc = s_ray[c - (x % stride) - (y % stride) * PW];
        const float ac = (nx * rc.x + ny * rc.y) + nz;
        accept = conf >= conf_thr && a >= 0.1f && ac >= 0.1f && inv_n >= 0.0f && inv_n <= 1.0f;
        if (accept) {
            // the plane is stored now; its cost (a full evaluation over all
            // sources, about 200 taps) is left to proposal_cost_kernel over
            // the compacted list, so that a wave does not idle on it
            const float inv = J.inv_min + (J.inv_max - J.inv_min) * inv_n;
            const float z = ((1.0f / inv) * ac) / a;
            ++ne;
            ++na;
            proposal = make_float4(nx, ny, nz, z);
            if (!prop) J.plane[0][i] = proposal;
        }
        // with `prop` the proposal is an extra hypothesis: every pixel runs
        // the plain sweep, and proposal_cost_kernel keeps the cheaper plane
        if (EV < 3 && (!accept || prop)) {
            unsigned ne2 = 0;
            init_tile<TW, TH, EV>(J, srcs, tl.y, tl.z, nh, s_cen, s_ray, ne2, s_tab);
            ne += ne2;
        }
    }
    // slots of the accepted pixels in the list: one global atomic per block
    unsigned slot = 0;
    if (accept) slot = atomicAdd(&s_acc, 1u);
    __syncthreads();
    if (threadIdx.x == 0) s_base = s_acc ? atomicAdd(accepted, (u64)s_acc) : 0;
    __syncthreads();
    if (accept) {
        list[s_base + slot] = ((u64)tl.x << 42) | ((u64)y << 21) | (u64)x;
        if (prop) prop[s_base + slot] = proposal;
    }
    if (EV >= 3) {
        // the pixels without an accepted proposal (all, with `prop`) sweep together
        unsigned ne2 = 0;
        init_box_tile<TW, TH>(J, srcs, tl.y, tl.z, nh, !accept || prop != nullptr, s_cen, s_ray, s_term, ne2, s_tab);
        ne += ne2;
    }
    (void)na;
    atomicAdd(&s_cnt, ne);
    __syncthreads();
    if (threadIdx.x == 0 && counter) atomicAdd(counter, (u64)s_cnt);
}

template <int TW, int TH, int EV>
__global__ __launch_bounds__(TW* TH) void proposal_init_kernel(const JobDesc* __restrict__ jobs,
                                                               const SrcDesc* __restrict__ srcs,
                                                               const int4* __restrict__ tiles, int ntiles, int nh,
                                                               const float* const* __restrict__ preds, int stride,
                                                               float conf_thr, u64* __restrict__ counter,
                                                               u64* __restrict__ accepted, u64* __restrict__ list,
                                                               float4* __restrict__ prop) {
    proposal_init_tile<TW, TH, EV>(jobs, srcs, tiles[xcd_remap(blockIdx.x, ntiles)], nh, preds, stride, conf_thr,
                                   counter, accepted, list, prop);
}

// The same over the 8 x 8 sub-tiles of a sorted launch's TW x TH tiles (one
// wave each, as init_sub_kernel): a smaller block stops costing a later
// source as soon as its own 64 pixels are done with it.  Sub-tile origins
// are multiples of 8, so a pixel's coarse pixel lies in its sub-tile.
template <int TW, int TH>
__global__ __launch_bounds__(64) void proposal_init_sub_kernel(const JobDesc* __restrict__ jobs,
                                                               const SrcDesc* __restrict__ srcs,
                                                               const int4* __restrict__ tiles, int ntiles, int nh,
                                                               const float* const* __restrict__ preds, int stride,
                                                               float conf_thr, u64* __restrict__ counter,
                                                               u64* __restrict__ accepted, u64* __restrict__ list,
                                                               float4* __restrict__ prop) {
    static_assert(TW % 8 == 0 && TH % 8 == 0, "whole 8x8 sub-tiles");
    constexpr int NSX = TW / 8, NSUB = NSX * (TH / 8);
    const int b = xcd_remap(blockIdx.x, ntiles * NSUB);
    const int sub = b % NSUB;
    const int4 tl = tiles[b / NSUB];
    const int x0 = tl.y + (sub % NSX) * 8, y0 = tl.z + (sub / NSX) * 8;
    if (x0 >= jobs[tl.x].w || y0 >= jobs[tl.x].h) return;   // the same for the whole block
    proposal_init_tile<8, 8, 3>(jobs, srcs, make_int4(tl.x, x0, y0, 0), nh, preds, stride, conf_thr, counter,
                                accepted, list, prop);
}

// Costs of the accepted proposals: one thread per entry of the compacted
// list (job << 42 | y << 21 | x), plane_cost over all sources with no abort,
// the reference loop order with clamped global window reads.  The list
// length is read from the device; a fixed grid strides over it.
// With `prop` (extra-hypothesis mode) the entry's plane comes from prop[k]
// and replaces the plain sweep's plane only when its cost is strictly lower:
// a 49th hypothesis, costed in full (the sweep's early abort does not apply
// to it, so a proposal whose partial mean once passed the sweep's threshold
// still competes with its final cost).
__global__ __launch_bounds__(256) void proposal_cost_kernel(const JobDesc* __restrict__ jobs,
                                                            const SrcDesc* __restrict__ srcs,
                                                            const u64* __restrict__ list,
                                                            const u64* __restrict__ count,
                                                            const float4* __restrict__ prop) {
    const u64 n = *count;
    for (u64 k = (u64)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (u64)gridDim.x * blockDim.x) {
        const u64 e = list[k];
        const int j = (int)(e >> 42), y = (int)((e >> 21) & 0x1FFFFF), x = (int)(e & 0x1FFFFF);
        const JobDesc& J = jobs[j];
        const size_t i = (size_t)y * J.w + x;
        const float4 pl = prop ? prop[k] : J.plane[0][i];
        const float2 ri = J.rays[i];
        const float a = (pl.x * ri.x + pl.y * ri.y) + pl.z;
        const GlobRef R{J.census, J.rays, x, y, J.w, J.h};
        const float dp = pl.w * a;
        const float cst = fminf(eval_plane<false>(R, srcs + J.src_off, J.nsrc, pl.x, pl.y, pl.z, dp,
                                                  __builtin_inff(), nullptr),
                                1.0f);
        if (!prop) {
            J.cost[0][i] = cst;
            J.dplane[i] = dp;
        } else if (cst < J.cost[0][i]) {
            J.plane[0][i] = pl;
            J.cost[0][i] = cst;
            J.dplane[i] = dp;
        }
    }
}

// Network input [n][Hp][Wp][64] bf16: channels 0-31 sweep features, 32 the
// normalised gray level, 33-63 zero; zero outside the h x w image.
__global__ void assemble_input_kernel(const uint16_t* const* __restrict__ vols, const uint8_t* const* __restrict__ grays,
                                      int n, int h, int w, int Hp, int Wp, int stride, int wfull,
                                      uint16_t* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;   // one thread per (pixel, 8 channels)
    const int64_t total = (int64_t)n * Hp * Wp * 8;
    if (t >= total) return;
    const int g = (int)(t & 7);
    const int64_t p = t >> 3;
    const int x = (int)(p % Wp);
    const int y = (int)((p / Wp) % Hp);
    const int b = (int)(p / ((int64_t)Wp * Hp));
    typedef unsigned short us8 __attribute__((ext_vector_type(8)));
    us8 v = us8{0, 0, 0, 0, 0, 0, 0, 0};
    if (x < w && y < h) {
        const size_t i = (size_t)y * w + x;
        if (g < 4) {
            v = *(const us8*)(vols[b] + i * 32 + 8 * g);
        } else if (g == 4) {
            const size_t gi = (size_t)(stride * y) * wfull + (size_t)(stride * x);
            v[0] = f2bf16((float)grays[b][gi] * (1.0f / 255.0f) - 0.5f);
        }
    }
    *(us8*)(out + (size_t)p * 64 + 8 * g) = v;
}

// Plain device copy (counter calibration: reads and writes n16 * 16 bytes).
__global__ void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void fill_u64_kernel(u64* p, int n, u64 v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// --------------------------------------------------------------------------
// instrumentation kernels
// --------------------------------------------------------------------------

__global__ void debug_cost_kernel(JobDesc J, const SrcDesc* __restrict__ S, const float* __restrict__ q, int n,
                                  float* cost, int* ham, int* ninv, float* peak) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float* e = q + 7 * t;
    const int px = (int)e[0], py = (int)e[1];
    const GlobRef R{J.census, J.rays, px, py, J.w, J.h};
    EvalInfo info;
    cost[t] = eval_plane<true>(R, S, J.nsrc, e[2], e[3], e[4], e[5], e[6] + 0.12f, &info);
    ham[t] = info.ham;
    ninv[t] = info.ninv;
    peak[t] = info.peak;
}

__global__ void debug_div_kernel(const float* a, const float* b, int n, float* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = div_rn_fast(a[i], b[i]);
}

__global__ void debug_cvt_kernel(const float* x, int n, int* out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = cvt_i32_rust(x[t]);
}

__global__ void debug_rng_kernel(const int* xy, const u64* salt, int n, int ndraw, u64* seed_out, float* uni) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    u64 s = pixel_rng(xy[2 * t], xy[2 * t + 1], salt[t]);
    seed_out[t] = s;
    for (int k = 0; k < ndraw; k++) uni[(size_t)t * ndraw + k] = uniform01(s);
}

// --------------------------------------------------------------------------
// launchers
// --------------------------------------------------------------------------

hipError_t launch_prepare(const PrepView* d_views, int nviews, int nblocks, hipStream_t st) {
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(prepare_kernel, dim3(nblocks), dim3(CT_W * CT_H), 0, st, d_views, nviews);
    return hipGetLastError();
}

int prepare_tiles(int w, int h, int& tx) {
    tx = (w + CT_W - 1) / CT_W;
    return tx * ((h + CT_H - 1) / CT_H);
}

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t st) {
    const size_t n16 = bytes / 16;
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, st, (const uint4*)src, (uint4*)dst, n16);
    return hipGetLastError();
}

hipError_t launch_fill_u64(u64* p, int n, u64 v, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_u64_kernel, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

#define TILE_CASES(X) \
    X(16, 16, 0) X(16, 16, 1) X(16, 16, 2) X(32, 8, 0) X(32, 8, 1) X(32, 8, 2) X(16, 8, 0) X(16, 8, 1) X(16, 8, 2) \
    X(8, 8, 1) X(8, 8, 2) X(16, 16, 3) X(16, 16, 4) X(16, 8, 3) X(16, 8, 4) X(8, 8, 3) X(8, 8, 4) X(32, 8, 3)

// sorted-form shapes (one wave per tile, evaluator 3)
#define SORTED_CASES(X) X(16, 8) X(24, 8) X(16, 16)

static hipError_t launch_init_sub(const TileCfg& cfg, const JobDesc* jobs, const SrcDesc* srcs, const int4* tiles,
                                  int ntiles, int nh, u64* counter, hipStream_t st) {
#define X(W, H)                                                                                                   \
    if (cfg.tw == W && cfg.th == H && cfg.eval == 3) {                                                            \
        hipLaunchKernelGGL((init_sub_kernel<W, H>), dim3(ntiles * (W / 8) * (H / 8)), dim3(64), 0, st, jobs, srcs, \
                           tiles, ntiles, nh, counter);                                                           \
        return hipGetLastError();                                                                                 \
    }
    SORTED_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_init(const TileCfg& cfg, const JobDesc* jobs, const SrcDesc* srcs, const int4* tiles, int ntiles,
                       int nh, u64* counter, hipStream_t st) {
    if (cfg.persistent == 2) return launch_init_sub(cfg, jobs, srcs, tiles, ntiles, nh, counter, st);
#define X(W, H, E)                                                                                     \
    if (cfg.tw == W && cfg.th == H && cfg.eval == E) {                                                 \
        hipLaunchKernelGGL((init_kernel<W, H, E>), dim3(ntiles), dim3(W * H), 0, st, jobs, srcs, tiles, ntiles, \
                           nh, counter);                                                               \
        return hipGetLastError();                                                                      \
    }
    TILE_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_pass(const TileCfg& cfg, const JobDesc* jobs, const SrcDesc* srcs, const int4* tiles, int ntiles,
                       int par, u64 salt, float sig_inv, float sig_n, u64* counter, int* queue, int grid_persistent,
                       hipStream_t st) {
    if (cfg.persistent == 2) {
#define X(W, H)                                                                                                  \
        if (cfg.tw == W && cfg.th == H && cfg.eval == 3) {                                                       \
            hipLaunchKernelGGL((pass_sorted_kernel<W, H, 3>), dim3(grid_persistent), dim3(64), 0, st, jobs, srcs, \
                               tiles, ntiles, par, salt, sig_inv, sig_n, counter, queue);                        \
            return hipGetLastError();                                                                            \
        }
        SORTED_CASES(X)
#undef X
        return hipErrorInvalidValue;
    }
#define X(W, H, E)                                                                                                 \
    if (cfg.tw == W && cfg.th == H && cfg.eval == E) {                                                             \
        if (cfg.persistent) {                                                                                      \
            hipLaunchKernelGGL((pass_persistent_kernel<W, H, E>), dim3(grid_persistent), dim3(W * H), 0, st, jobs, \
                               srcs, tiles, ntiles, par, salt, sig_inv, sig_n, counter, queue);                    \
        } else {                                                                                                   \
            hipLaunchKernelGGL((pass_kernel<W, H, E>), dim3(ntiles), dim3(W * H), 0, st, jobs, srcs, tiles, ntiles, \
                               par, salt, sig_inv, sig_n, counter);                                                \
        }                                                                                                          \
        return hipGetLastError();                                                                                  \
    }
    TILE_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_gate(const TileCfg& cfg, const JobDesc* jobs, const int4* tiles, int ntiles, int par,
                       float cost_thr, hipStream_t st) {
#define X(W, H)                                                                                                  \
    if (cfg.tw == W && cfg.th == H) {                                                                            \
        hipLaunchKernelGGL((gate_kernel<W, H>), dim3(ntiles), dim3(W * H), 0, st, jobs, tiles, par, cost_thr);    \
        return hipGetLastError();                                                                                \
    }
    X(16, 16) X(32, 8) X(16, 8) X(8, 8) X(24, 8)
#undef X
    return hipErrorInvalidValue;
}

int persistent_blocks_per_cu(const TileCfg& cfg) {
    int nb = 0;
#define X(W, H)                                                                                          \
    if (cfg.persistent == 2 && cfg.tw == W && cfg.th == H && cfg.eval == 3) {                            \
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pass_sorted_kernel<W, H, 3>, 64, 0);    \
        return nb > 0 ? nb : 1;                                                                          \
    }
    SORTED_CASES(X)
#undef X
#define X(W, H, E)                                                                                           \
    if (cfg.tw == W && cfg.th == H && cfg.eval == E) {                                                       \
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pass_persistent_kernel<W, H, E>, W * H, 0); \
    }
    TILE_CASES(X)
#undef X
    return nb > 0 ? nb : 1;
}

hipError_t launch_volume(const JobDesc* jobs, const SrcDesc* srcs, const int4* tiles, int ntiles, int nplanes,
                         int stride, const VolOut* outs, hipStream_t st) {
    if (stride == 1)
        hipLaunchKernelGGL((volume_kernel<16, 16, 1>), dim3(ntiles), dim3(256), 0, st, jobs, srcs, tiles, ntiles, nplanes,
                           outs);
    else if (stride == 2)
        hipLaunchKernelGGL((volume_kernel<16, 16, 2>), dim3(ntiles), dim3(256), 0, st, jobs, srcs, tiles, ntiles, nplanes,
                           outs);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

#define PROPOSAL_CASES(X) \
    X(16, 16, 0) X(16, 16, 1) X(16, 16, 2) X(16, 16, 3) X(16, 16, 4) X(16, 8, 3) X(16, 8, 4) X(8, 8, 3) X(8, 8, 4) \
    X(32, 8, 3) X(24, 8, 3)

bool proposal_supported(const TileCfg& cfg) {
#define X(W, H, E) \
    if (cfg.tw == W && cfg.th == H && cfg.eval == E) return true;
    PROPOSAL_CASES(X)
#undef X
    return false;
}

hipError_t launch_proposal_init(const TileCfg& cfg, const JobDesc* jobs, const SrcDesc* srcs, const int4* tiles,
                                int ntiles, int nh, const float* const* preds, int stride, float conf_thr,
                                u64* counter, u64* accepted, u64* list, float4* prop, hipStream_t st) {
    bool done = false;
    if (cfg.persistent == 2) {
#define X(W, H)                                                                                                   \
        if (!done && cfg.tw == W && cfg.th == H && cfg.eval == 3) {                                               \
            hipLaunchKernelGGL((proposal_init_sub_kernel<W, H>), dim3(ntiles * (W / 8) * (H / 8)), dim3(64), 0, st, \
                               jobs, srcs, tiles, ntiles, nh, preds, stride, conf_thr, counter, accepted, list,     \
                               prop);                                                                           \
            done = true;                                                                                        \
        }
        SORTED_CASES(X)
#undef X
    }
#define X(W, H, E)                                                                                              \
    if (!done && cfg.tw == W && cfg.th == H && cfg.eval == E) {                                                 \
        hipLaunchKernelGGL((proposal_init_kernel<W, H, E>), dim3(ntiles), dim3(W * H), 0, st, jobs, srcs, tiles, \
                           ntiles, nh, preds, stride, conf_thr, counter, accepted, list, prop);                 \
        done = true;                                                                                            \
    }
    PROPOSAL_CASES(X)
#undef X
    if (!done) return hipErrorInvalidValue;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(proposal_cost_kernel, dim3(2048), dim3(256), 0, st, jobs, srcs, list, accepted,
                       (const float4*)prop);
    return hipGetLastError();
}

hipError_t launch_assemble(const uint16_t* const* vols, const uint8_t* const* grays, int n, int h, int w, int Hp,
                           int Wp, int stride, int wfull, uint16_t* out, hipStream_t st) {
    const int64_t total = (int64_t)n * Hp * Wp * 8;
    hipLaunchKernelGGL(assemble_input_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, vols, grays, n,
                       h, w, Hp, Wp, stride, wfull, out);
    return hipGetLastError();
}

hipError_t launch_debug_cost(const JobDesc& J, const SrcDesc* S, const float* q, int n, float* cost, int* ham,
                             int* ninv, float* peak, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(debug_cost_kernel, dim3((n + 63) / 64), dim3(64), 0, st, J, S, q, n, cost, ham, ninv, peak);
    return hipGetLastError();
}

hipError_t launch_debug_div(const float* a, const float* b, int n, float* out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(debug_div_kernel, dim3((n + 255) / 256), dim3(256), 0, st, a, b, n, out);
    return hipGetLastError();
}

hipError_t launch_debug_cvt(const float* x, int n, int* out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(debug_cvt_kernel, dim3((n + 63) / 64), dim3(64), 0, st, x, n, out);
    return hipGetLastError();
}

hipError_t launch_debug_rng(const int* xy, const u64* salt, int n, int ndraw, u64* seed_out, float* uni,
                            hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(debug_rng_kernel, dim3((n + 63) / 64), dim3(64), 0, st, xy, salt, n, ndraw, seed_out, uni);
    return hipGetLastError();
}

}  // namespace mvs
