// image.hip -- input stage of load_views (mvs/src/lib.rs:185-197) on the device:
// the working-resolution resize of a decoded u8 image and its luma.
//
// The reference calls the `image` crate (0.24.9): resize_exact with the
// Triangle filter -- a vertical pass into f32, then a horizontal pass -- and
// to_luma8 / to_rgb8.  Per output row (column): ratio = in / out (f32),
// sratio = max(ratio, 1), taps left = clamp(floor(centre - sratio), 0, in-1)
// .. right = clamp(ceil(centre + sratio), left+1, in) around
// centre = (o + 0.5) * ratio, weights max(1 - |(i - (centre - 0.5)) / sratio|, 0)
// normalised by their f32 sum, values accumulated in f32 in tap order; the
// horizontal pass clamps to [0, 255] and rounds half away from zero.  Luma
// of RGB is rgb_to_luma of the crate's color.rs: (2126 r + 7152 g + 722 b)
// / 10000 in u32 with truncation.  One thread per output pixel and pass;
// all channels of a pixel in the same thread.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace mvs {

__device__ __forceinline__ void tri_window(int o, int n_in, int n_out, int& left, int& right, float& c, float& sratio) {
    const float ratio = (float)n_in / (float)n_out;
    sratio = ratio < 1.0f ? 1.0f : ratio;
    const float support = 1.0f * sratio;
    const float centre = ((float)o + 0.5f) * ratio;
    int l = (int)floorf(centre - support);
    l = min(max(l, 0), n_in - 1);
    int r = (int)ceilf(centre + support);
    r = min(max(r, l + 1), n_in);
    left = l;
    right = r;
    c = centre - 0.5f;
}

__device__ __forceinline__ float tri_weight(int i, float c, float sratio) {
    const float ax = fabsf(((float)i - c) / sratio);
    return ax < 1.0f ? 1.0f - ax : 0.0f;
}

// src u8 [h][w][C] -> tmp f32 [mh][w][C]
template <int C>
__global__ void resize_vertical_kernel(const uint8_t* __restrict__ src, int w, int h, int mh, float* __restrict__ tmp) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, oy = blockIdx.y;
    if (x >= w) return;
    int l, r;
    float c, sr;
    tri_window(oy, h, mh, l, r, c, sr);
    float s = 0.0f;
    for (int i = l; i < r; ++i) s += tri_weight(i, c, sr);
    float t[C];
    for (int k = 0; k < C; ++k) t[k] = 0.0f;
    for (int i = l; i < r; ++i) {
        const float wt = tri_weight(i, c, sr) / s;
        const uint8_t* p = src + ((size_t)i * w + x) * C;
        for (int k = 0; k < C; ++k) t[k] += (float)p[k] * wt;
    }
    float* o = tmp + ((size_t)oy * w + x) * C;
    for (int k = 0; k < C; ++k) o[k] = t[k];
}

// tmp f32 [mh][w][C] -> out u8 [mh][mw][C]
template <int C>
__global__ void resize_horizontal_kernel(const float* __restrict__ tmp, int w, int mw, int mh, uint8_t* __restrict__ out) {
    const int ox = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (ox >= mw || y >= mh) return;
    int l, r;
    float c, sr;
    tri_window(ox, w, mw, l, r, c, sr);
    float s = 0.0f;
    for (int i = l; i < r; ++i) s += tri_weight(i, c, sr);
    float t[C];
    for (int k = 0; k < C; ++k) t[k] = 0.0f;
    for (int i = l; i < r; ++i) {
        const float wt = tri_weight(i, c, sr) / s;
        const float* p = tmp + ((size_t)y * w + i) * C;
        for (int k = 0; k < C; ++k) t[k] += p[k] * wt;
    }
    uint8_t* o = out + ((size_t)y * mw + ox) * C;
    for (int k = 0; k < C; ++k) o[k] = (uint8_t)roundf(fminf(fmaxf(t[k], 0.0f), 255.0f));
}

__global__ void luma_kernel(const uint8_t* __restrict__ rgb, int64_t n, uint8_t* __restrict__ gray) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t l = 2126u * rgb[3 * i] + 7152u * rgb[3 * i + 1] + 722u * rgb[3 * i + 2];
    gray[i] = (uint8_t)(l / 10000u);   // <= 255: the sum of the weights is the divisor
}

hipError_t launch_resize(const uint8_t* src, int channels, int w, int h, int mw, int mh, float* tmp, uint8_t* out,
                         hipStream_t st) {
    const dim3 b(128);
    if (channels == 1) {
        hipLaunchKernelGGL(resize_vertical_kernel<1>, dim3((w + 127) / 128, mh), b, 0, st, src, w, h, mh, tmp);
        hipLaunchKernelGGL(resize_horizontal_kernel<1>, dim3((mw + 127) / 128, mh), b, 0, st, tmp, w, mw, mh, out);
    } else if (channels == 3) {
        hipLaunchKernelGGL(resize_vertical_kernel<3>, dim3((w + 127) / 128, mh), b, 0, st, src, w, h, mh, tmp);
        hipLaunchKernelGGL(resize_horizontal_kernel<3>, dim3((mw + 127) / 128, mh), b, 0, st, tmp, w, mw, mh, out);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_luma(const uint8_t* rgb, int64_t n, uint8_t* gray, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(luma_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, rgb, n, gray);
    return hipGetLastError();
}

}  // namespace mvs
