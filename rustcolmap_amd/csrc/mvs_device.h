// mvs_device.h -- device-side data layout and exact arithmetic helpers shared
// by the stereo and fusion kernels (gfx950, wave64).
//
// Every f32/f64 expression below keeps the reference's operation order
// (mvs/src/stereo.rs, mvs/src/fuse.rs); the library is compiled with
// -ffp-contract=off and the default correctly-rounded f32 divide / sqrt, so
// each operation rounds exactly as the reference's IEEE arithmetic does.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace mvs {

typedef unsigned long long u64;

// Per (job, source) warp parameters: SourceCtx (stereo.rs:84-96).
struct SrcDesc {
    const u64* census;
    int32_t w, h;
    float rs[9];
    float ts[3];
    float f, cx, cy, k;
};

// One reference view of a stereo batch.
struct JobDesc {
    const u64* census;      // reference census
    const float2* rays;     // (rx, ry) per pixel (RefWorkspace::new, stereo.rs:115-125)
    float4* plane[2];       // ping-pong (nx, ny, nz, z)
    float* cost[2];         // ping-pong cost
    float* depth_out;       // gated depth (0 = invalid)
    float* cost_out;
    float* normal_out;      // 3 floats per pixel
    int32_t w, h;
    int32_t nsrc, src_off;  // sources at srcs[src_off .. src_off + nsrc)
    float zmin, zmax;       // f32 as RefWorkspace stores them (stereo.rs:150)
    float inv_min, inv_max; // 1/zmax, 1/zmin (stereo.rs:248-249)
    int32_t tiles_x, tiles_y;
    int32_t tile_off;       // first tile of this job in the batch tile space
    int32_t pinhole_ok;     // k == 0 sources may skip the distortion polynomial (finite-r2 bound holds)
    int32_t costs_exact;    // every stored cost is at most its plane's cost here (states made by the init
                            // sweep and the passes, not planted ones): the pass may then skip a neighbour
                            // candidate equal to the current plane
    float* dplane;          // per pixel: the d_plane the stored plane was costed with, written wherever a
                            // plane other than (+0, +0, 1, z) is stored (for those d_plane = z); one buffer,
                            // each pixel reading and writing only its own entry
    u64* valid_count;       // gate: number of valid pixels (accumulated), or NULL
    uint8_t* work;          // per pixel: window rows read in the last pass / 2 (capped at 255), or NULL;
                            // orders the pixels of a tile in the sorted pass kernel, never read for results
};

// A view as the fusion kernels see it (fuse.rs:99-162).
struct FuseView {
    const float* depth;     // gated depth, NULL = no depth map
    const uint8_t* rgb;     // 3 per pixel, or NULL = gray replicated
    const uint8_t* gray;
    int32_t w, h;
    double f, cx, cy, k;
    double R[9];
    double t[3];
    int64_t pix_off;        // first global candidate index of this view
    int32_t ord_off, ord_n; // fusion_order slice
};

// --------------------------------------------------------------------------
// Rust `as` casts
// --------------------------------------------------------------------------

// Validity of `(x as i32)` against [0, n): Rust truncates toward zero,
// saturates and maps NaN to 0, so the integer lands in [0, n) iff x is NaN
// (-> 0, valid for n > 0) or -1 < x < n.  Returns the index when valid.
__device__ __forceinline__ bool cast_in_range(float x, int32_t n, int32_t& out) {
    if (x != x) { out = 0; return n > 0; }
    if (x > -1.0f && x < (float)n) { out = (int32_t)x; return true; }
    return false;
}

__device__ __forceinline__ int32_t f64_to_i32_sat(double x) {
    if (x != x) return 0;
    if (x >= 2147483648.0) return 2147483647;
    if (x <= -2147483648.0) return (-2147483647 - 1);
    return (int32_t)x;
}

__device__ __forceinline__ long long f64_to_i64_sat(double x) {
    if (x != x) return 0;
    if (x >= 9223372036854775808.0) return 9223372036854775807LL;
    if (x <= -9223372036854775808.0) return (-9223372036854775807LL - 1);
    return (long long)x;
}

__device__ __forceinline__ uint8_t f64_to_u8_sat(double x) {
    if (x != x || x <= 0.0) return 0;
    if (x >= 255.0) return 255;
    return (uint8_t)x;
}

// popcount / 48 rounded to f32, as `count_ones() as f32 / BITS`
// (stereo.rs:213).  q = pc * RN(1/48) corrected by one FMA residual step
// equals the correctly rounded quotient for every pc in 0..=48 (checked
// exhaustively by the test suite).
__device__ __forceinline__ float pc_over_48(int pc) {
    const float r = 1.0f / 48.0f;
    const float a = (float)pc;
    const float q = a * r;
    const float rem = __builtin_fmaf(-q, 48.0f, a);
    return __builtin_fmaf(rem, r, q);
}

// --------------------------------------------------------------------------
// RNG (stereo.rs:47-78)
// --------------------------------------------------------------------------

__device__ __forceinline__ u64 pixel_rng(int32_t x, int32_t y, u64 salt) {
    u64 s = ((u64)(long long)x * 0x9E3779B97F4A7C15ULL) ^ ((u64)(long long)y * 0xC2B2AE3D27D4EB4FULL) ^
            (salt * 0x165667B19E3779F9ULL);
    s ^= s >> 30;
    s *= 0xBF58476D1CE4E5B9ULL;
    s ^= s >> 27;
    s *= 0x94D049BB133111EBULL;
    return s ^ (s >> 31);
}

__device__ __forceinline__ u64 next_u64(u64& s) {
    u64 x = s;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    s = x;
    return x;
}

// (x >> 40) as f32 / 2^24: a 24-bit integer scaled by a power of two, exact.
__device__ __forceinline__ float uniform01(u64& s) {
    return (float)(uint32_t)(next_u64(s) >> 40) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ float gauss3(u64& s) {
    const float u1 = uniform01(s);
    const float u2 = uniform01(s);
    const float u3 = uniform01(s);
    return ((u1 + u2) + u3) - 1.5f;
}

// --------------------------------------------------------------------------
// geometry (sfm-lite/src/geom.rs:27-42, 63-76), f64
// --------------------------------------------------------------------------

__device__ __forceinline__ void cam_normalize(double f, double cx, double cy, double k, double u, double v,
                                              double& xo, double& yo) {
    double xn = (u - cx) / f;
    double yn = (v - cy) / f;
    for (int it = 0; it < 2; it++) {
        const double r2 = xn * xn + yn * yn;
        const double d = 1.0 + k * r2;
        if (fabs(d) < 1e-12) break;
        xn = (u - cx) / (f * d);
        yn = (v - cy) / (f * d);
    }
    xo = xn;
    yo = yn;
}

// R x (+ t), nalgebra's left-to-right row dot products.
__device__ __forceinline__ void mat_vec(const double* R, const double* x, double* o) {
    o[0] = (R[0] * x[0] + R[1] * x[1]) + R[2] * x[2];
    o[1] = (R[3] * x[0] + R[4] * x[1]) + R[5] * x[2];
    o[2] = (R[6] * x[0] + R[7] * x[1]) + R[8] * x[2];
}
// R^T x
__device__ __forceinline__ void mat_t_vec(const double* R, const double* x, double* o) {
    o[0] = (R[0] * x[0] + R[3] * x[1]) + R[6] * x[2];
    o[1] = (R[1] * x[0] + R[4] * x[1]) + R[7] * x[2];
    o[2] = (R[2] * x[0] + R[5] * x[1]) + R[8] * x[2];
}

}  // namespace mvs
