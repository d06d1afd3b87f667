// net_kernels.hip -- kernels of the learned plane-proposal network (gfx950).
//
// A 5-level U-Net of 3x3 convolutions maps the reference image plus a
// 32-plane census sweep volume to per-pixel normalised inverse depth, unit
// normal and confidence.  Activations are NHWC bf16, weights bf16 with fp32
// master copies, accumulation fp32 on MFMA (v_mfma_f32_16x16x32_bf16).
//
// conv3x3_kernel is an implicit GEMM: M = output pixels (a 16x8 tile per
// workgroup, 32 per wave), N = output channels (32 per workgroup), K = 9
// taps x input channels in chunks of 32.  Input fragments are 16-byte loads
// of 8 consecutive channels of one pixel (NHWC), weight fragments 16-byte
// loads from a [tap][cout][cin] copy.  The same kernel computes the data
// gradient from a flipped [tap][cin][cout] copy.  The input may be the
// concatenation of a full-resolution tensor and a nearest-upsampled
// half-resolution tensor (the U-Net skip join), read in place.
//
// wgrad_kernel: dW[tap][ci][co] = sum_p X[p + tap][ci] dY[p][co]; tiles of
// 32 pixels of X and dY are staged transposed in LDS so both MFMA operands
// are contiguous; partial sums over pixel chunks are added atomically.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "net_kernels.h"

namespace mvs {
namespace net {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ bf16x8 as_bf(us8 v) { return __builtin_bit_cast(bf16x8, v); }

// --------------------------------------------------------------------------
// conv3x3 (forward and data gradient)
// --------------------------------------------------------------------------

__global__ __launch_bounds__(256) void conv3x3_kernel(ConvArgs a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tiles_x = (a.W + 15) >> 4;
    const int tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
    const int co0 = blockIdx.y * 32;
    const int n = blockIdx.z;
    const int cin = a.c0 + a.c1;
    const int r16 = lane & 15, kq = lane >> 4;     // fragment row / k-group
    // the two 16-pixel fragments of this wave: rows 2*wave and 2*wave+1 of the 16x8 tile
    int px[2], py[2];
    for (int mf = 0; mf < 2; ++mf) {
        px[mf] = tx * 16 + r16;
        py[mf] = ty * 8 + 2 * wave + mf;
    }
    f32x4 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int t = 0; t < 9; ++t) {
        const int dy = t / 3 - 1, dx = t % 3 - 1;
        for (int ci0 = 0; ci0 < cin; ci0 += 32) {
            us8 av[2];
            for (int mf = 0; mf < 2; ++mf) {
                const int ix = px[mf] + dx, iy = py[mf] + dy;
                us8 v = us8{0, 0, 0, 0, 0, 0, 0, 0};
                if (ix >= 0 && iy >= 0 && ix < a.W && iy < a.H) {
                    const int c = ci0 + 8 * kq;
                    const uint16_t* src;
                    if (ci0 < a.c0)
                        src = a.x0 + (((size_t)n * a.H + iy) * a.W + ix) * a.c0 + c;
                    else
                        src = a.x1 + (((size_t)n * a.H1 + (iy >> 1)) * a.W1 + (ix >> 1)) * a.c1 + (c - a.c0);
                    v = *(const us8*)src;
                }
                av[mf] = v;
            }
            us8 bv[2];
            for (int nf = 0; nf < 2; ++nf) {
                const int co = co0 + 16 * nf + r16;
                bv[nf] = *(const us8*)(a.wt + ((size_t)t * a.cout + co) * cin + ci0 + 8 * kq);
            }
            for (int mf = 0; mf < 2; ++mf)
                for (int nf = 0; nf < 2; ++nf)
                    acc[mf][nf] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(av[mf]), as_bf(bv[nf]), acc[mf][nf],
                                                                         0, 0, 0);
        }
    }
    // D[m][n]: lane holds rows m = 4*kq + i, column n = r16
    for (int mf = 0; mf < 2; ++mf) {
        for (int nf = 0; nf < 2; ++nf) {
            const int co = co0 + 16 * nf + r16;
            const float b = a.bias ? a.bias[co] : 0.0f;
            for (int i = 0; i < 4; ++i) {
                const int m = 4 * kq + i;   // pixel within the 16-pixel fragment
                const int ox = tx * 16 + m, oy = ty * 8 + 2 * wave + mf;
                if (ox >= a.W || oy >= a.H) continue;
                float v = acc[mf][nf][i] + b;
                const size_t pix = ((size_t)n * a.H + oy) * a.W + ox;
                if (a.mode == 0) {
                    if (a.relu) v = fmaxf(v, 0.0f);
                    a.y[pix * a.cout + co] = f2bf(v);
                } else {
                    if (co < a.split) {
                        a.yf0[pix * a.split + co] += v;
                    } else {
                        const size_t hp = ((size_t)n * a.H1 + (oy >> 1)) * a.W1 + (ox >> 1);
                        atomicAdd(&a.yf1[hp * (a.cout - a.split) + (co - a.split)], v);
                    }
                }
            }
        }
    }
}

// Forward convolution with LDS staging (mode 0): a workgroup computes a
// 16x16-pixel x 32-channel output tile; per 32-channel input chunk it stages
// the 18x18 input halo tile ([pixel][32 ch], 64 B per pixel) and the 9 taps'
// weights ([tap][co][32 ci]) in LDS once, then each wave runs 9 taps x 4
// pixel fragments (its 4 rows) x 2 channel fragments of MFMA from 16-byte
// LDS reads -- 64 consecutive lanes read 1 KB of consecutive LDS, free of
// bank conflicts.  Same values as conv3x3_kernel in mode 0: each output is
// the fp32 MFMA sum over taps and channel chunks in the same order.
constexpr int CT = 16;              // output tile side
constexpr int CH = CT + 2;          // staged side with the halo

__global__ __launch_bounds__(256) void conv3x3_lds_kernel(ConvArgs a) {
    __shared__ __attribute__((aligned(16))) uint16_t s_in[CH * CH * 32];
    __shared__ __attribute__((aligned(16))) uint16_t s_w[9 * 32 * 32];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tiles_x = (a.W + CT - 1) / CT;
    const int tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
    const int co0 = blockIdx.y * 32;
    const int n = blockIdx.z;
    const int cin = a.c0 + a.c1;
    const int r16 = lane & 15, kq = lane >> 4;
    const int x0 = tx * CT - 1, y0 = ty * CT - 1;   // staged origin (halo)
    f32x4 acc[4][2];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int ci0 = 0; ci0 < cin; ci0 += 32) {
        __syncthreads();   // previous chunk's LDS reads are done
        // input halo tile: CH*CH pixels x 4 chunks of 8 channels
        for (int j = threadIdx.x; j < CH * CH * 4; j += 256) {
            const int pix = j >> 2, q = j & 3;
            const int ix = x0 + pix % CH, iy = y0 + pix / CH;
            us8 v = us8{0, 0, 0, 0, 0, 0, 0, 0};
            if (ix >= 0 && iy >= 0 && ix < a.W && iy < a.H) {
                const int c = ci0 + 8 * q;
                const uint16_t* src;
                if (ci0 < a.c0)
                    src = a.x0 + (((size_t)n * a.H + iy) * a.W + ix) * a.c0 + c;
                else
                    src = a.x1 + (((size_t)n * a.H1 + (iy >> 1)) * a.W1 + (ix >> 1)) * a.c1 + (c - a.c0);
                v = *(const us8*)src;
            }
            *(us8*)(s_in + pix * 32 + 8 * q) = v;
        }
        // weights of the 32 output channels: 9 taps x 32 co x 4 chunks
        for (int j = threadIdx.x; j < 9 * 32 * 4; j += 256) {
            const int q = j & 3, co = (j >> 2) & 31, t = j >> 7;
            *(us8*)(s_w + (t * 32 + co) * 32 + 8 * q) =
                *(const us8*)(a.wt + ((size_t)t * a.cout + co0 + co) * cin + ci0 + 8 * q);
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int dy = t / 3, dx = t % 3;   // offsets into the halo tile
            us8 bv[2];
#pragma unroll
            for (int nf = 0; nf < 2; ++nf) bv[nf] = *(const us8*)(s_w + (t * 32 + 16 * nf + r16) * 32 + 8 * kq);
#pragma unroll
            for (int mf = 0; mf < 4; ++mf) {
                const int row = 4 * wave + mf + dy, col = r16 + dx;
                const us8 av = *(const us8*)(s_in + (row * CH + col) * 32 + 8 * kq);
#pragma unroll
                for (int nf = 0; nf < 2; ++nf)
                    acc[mf][nf] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(av), as_bf(bv[nf]), acc[mf][nf], 0, 0,
                                                                         0);
            }
        }
    }
    // D[m][n]: lane holds rows m = 4*kq + i (pixel within the fragment), column n = r16
    for (int mf = 0; mf < 4; ++mf) {
        const int oy = ty * CT + 4 * wave + mf;
        if (oy >= a.H) continue;
        for (int nf = 0; nf < 2; ++nf) {
            const int co = co0 + 16 * nf + r16;
            const float b = a.bias ? a.bias[co] : 0.0f;
            for (int i = 0; i < 4; ++i) {
                const int ox = tx * CT + 4 * kq + i;
                if (ox >= a.W) continue;
                float v = acc[mf][nf][i] + b;
                if (a.relu) v = fmaxf(v, 0.0f);
                a.y[(((size_t)n * a.H + oy) * a.W + ox) * a.cout + co] = f2bf(v);
            }
        }
    }
}

// The fragment index m of an MFMA row within a 16-pixel fragment maps to
// pixel px = tx*16 + r16 on the A side; on the D side the row m is the
// pixel index.  Both describe the same 16 pixels of one tile row.

// --------------------------------------------------------------------------
// weight gradient
// --------------------------------------------------------------------------

__global__ __launch_bounds__(256) void wgrad_kernel(WgradArgs a) {
    __shared__ __attribute__((aligned(16))) uint16_t sx[32][40];    // [ci][px], padded rows (80 B)
    __shared__ __attribute__((aligned(16))) uint16_t sd[32][40];    // [co][px]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, kq = lane >> 4;
    const int nci = (a.c0 + a.c1) / 32;
    const int ci0 = (blockIdx.y % nci) * 32, co0 = (blockIdx.y / nci) * 32;
    const int t = blockIdx.z, dy = t / 3 - 1, dx = t % 3 - 1;
    const int64_t npix = (int64_t)a.N * a.H * a.W;
    const int64_t p_begin = (int64_t)blockIdx.x * a.chunk;
    const int64_t p_end = p_begin + a.chunk < npix ? p_begin + a.chunk : npix;
    const int mi = wave & 1, ni = wave >> 1;
    f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
    // loader mapping: 256 threads = 32 pixels x 8 groups; groups 0-3 load X
    // channels, groups 4-7 load dY channels, 8 channels each
    const int lp = threadIdx.x & 31, lg = threadIdx.x >> 5;
    for (int64_t p0 = p_begin; p0 < p_end; p0 += 32) {
        const int64_t p = p0 + lp;
        us8 v = us8{0, 0, 0, 0, 0, 0, 0, 0};
        if (p < p_end) {
            const int x = (int)(p % a.W);
            const int64_t r = p / a.W;
            const int y = (int)(r % a.H);
            const int nimg = (int)(r / a.H);
            if (lg < 4) {
                const int ix = x + dx, iy = y + dy;
                if (ix >= 0 && iy >= 0 && ix < a.W && iy < a.H) {
                    const int c = ci0 + 8 * lg;
                    const uint16_t* src;
                    if (ci0 < a.c0)
                        src = a.x0 + (((size_t)nimg * a.H + iy) * a.W + ix) * a.c0 + c;
                    else
                        src = a.x1 + (((size_t)nimg * a.H1 + (iy >> 1)) * a.W1 + (ix >> 1)) * a.c1 + (c - a.c0);
                    v = *(const us8*)src;
                }
            } else {
                v = *(const us8*)(a.dy + (size_t)p * a.cout + co0 + 8 * (lg - 4));
            }
        }
        if (lg < 4) {
            for (int k = 0; k < 8; ++k) sx[8 * lg + k][lp] = v[k];
        } else {
            for (int k = 0; k < 8; ++k) sd[8 * (lg - 4) + k][lp] = v[k];
        }
        __syncthreads();
        // A = X^T tile (16 ci x 32 px), B = dY tile (32 px x 16 co)
        const us8 av = *(const us8*)&sx[16 * mi + r16][8 * kq];
        const us8 bv = *(const us8*)&sd[16 * ni + r16][8 * kq];
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(av), as_bf(bv), acc, 0, 0, 0);
        __syncthreads();
    }
    const int cin = a.c0 + a.c1;
    for (int i = 0; i < 4; ++i) {
        const int ci = ci0 + 16 * mi + 4 * kq + i;
        const int co = co0 + 16 * ni + r16;
        atomicAdd(&a.dw[((size_t)t * cin + ci) * a.cout + co], acc[i]);
    }
}

// db[co] = sum_p dY[p][co]
__global__ __launch_bounds__(256) void bgrad_kernel(const uint16_t* __restrict__ dy, int64_t npix, int cout,
                                                    int64_t chunk, float* __restrict__ db) {
    const int c = threadIdx.x % cout;
    const int lanes = 256 / cout;       // cout divides 256? handled: cout in {32,64,96,128,160}
    const int sub = threadIdx.x / cout;
    if (sub >= lanes) return;
    const int64_t p0 = (int64_t)blockIdx.x * chunk;
    const int64_t p1 = p0 + chunk < npix ? p0 + chunk : npix;
    float s = 0.0f;
    for (int64_t p = p0 + sub; p < p1; p += lanes) s += bf2f(dy[(size_t)p * cout + c]);
    atomicAdd(&db[c], s);
}

// dY' = bf16(y > 0 ? dY : 0)   (ReLU backward, y is the post-ReLU output)
__global__ void relu_bwd_kernel(const float* __restrict__ g, const uint16_t* __restrict__ y, int64_t n,
                                uint16_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = (y[i] & 0x7fffu) && !(y[i] & 0x8000u) ? f2bf(g[i]) : (uint16_t)0;
}

// 2x2 average pooling, NHWC, H and W even
__global__ void pool_fwd_kernel(const uint16_t* __restrict__ x, int N, int H, int W, int C, uint16_t* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int H2 = H / 2, W2 = W / 2;
    const int64_t total = (int64_t)N * H2 * W2 * C;
    if (i >= total) return;
    const int c = (int)(i % C);
    int64_t r = i / C;
    const int x2 = (int)(r % W2);
    r /= W2;
    const int y2 = (int)(r % H2);
    const int n = (int)(r / H2);
    const size_t b = (((size_t)n * H + 2 * y2) * W + 2 * x2) * C + c;
    const float s = bf2f(x[b]) + bf2f(x[b + C]) + bf2f(x[b + (size_t)W * C]) + bf2f(x[b + (size_t)W * C + C]);
    y[i] = f2bf(0.25f * s);
}

__global__ void pool_bwd_kernel(const float* __restrict__ gy, int N, int H, int W, int C, float* __restrict__ gx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)N * H * W * C;
    if (i >= total) return;
    const int c = (int)(i % C);
    int64_t r = i / C;
    const int x = (int)(r % W);
    r /= W;
    const int y = (int)(r % H);
    const int n = (int)(r / H);
    gx[i] += 0.25f * gy[(((size_t)n * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c];
}

__global__ void zero_f32_kernel(float* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// --------------------------------------------------------------------------
// head: 1x1 conv 32 -> 5, predictions, loss and its gradient
// --------------------------------------------------------------------------

__device__ __forceinline__ float sigm(float x) { return 1.0f / (1.0f + __expf(-x)); }

__global__ __launch_bounds__(256) void head_kernel(HeadArgs a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ float s_w[5 * 32 + 5];
    __shared__ float s_red[8][5 * 32 + 5 + 4];
    for (int i = threadIdx.x; i < 5 * 32 + 5; i += blockDim.x) s_w[i] = a.w[i];
    for (int i = threadIdx.x; i < 8 * (5 * 32 + 9); i += blockDim.x) (&s_red[0][0])[i] = 0.0f;
    __syncthreads();
    const bool act = p < a.npix;
    float f[32];
    float o[5] = {0, 0, 0, 0, 0};
    int x = 0, y = 0, n = 0;
    bool inside = false;
    if (act) {
        x = (int)(p % a.W);
        y = (int)((p / a.W) % a.H);
        n = (int)(p / ((int64_t)a.W * a.H));
        inside = x < a.w_valid && y < a.h_valid;
        const uint16_t* fp = a.feat + (size_t)p * 32;
        for (int c = 0; c < 32; ++c) f[c] = bf2f(fp[c]);
        for (int k = 0; k < 5; ++k) {
            float s = s_w[5 * 32 + k];
            for (int c = 0; c < 32; ++c) s += s_w[k * 32 + c] * f[c];
            o[k] = s;
        }
        const float inv = sigm(o[0]);
        float nx = o[1], ny = o[2], nz = o[3];
        const float len = sqrtf(nx * nx + ny * ny + nz * nz) + 1e-6f;
        nx /= len;
        ny /= len;
        nz /= len;
        const float conf = sigm(o[4]);
        if (a.pred && inside) {
            const size_t q = ((size_t)n * a.h_valid + y) * a.w_valid + x;
            a.pred[5 * q + 0] = inv;
            a.pred[5 * q + 1] = nx;
            a.pred[5 * q + 2] = ny;
            a.pred[5 * q + 3] = nz;
            a.pred[5 * q + 4] = conf;
        }
        if (a.gfeat) {
            float go[5] = {0, 0, 0, 0, 0};
            float ld = 0.f, ln = 0.f, lc = 0.f;
            if (inside) {
                const size_t q = ((size_t)n * a.h_valid + y) * a.w_valid + x;
                const float* tg = a.target + 5 * q;   // inv, nx, ny, nz, valid
                const bool v = tg[4] > 0.5f;
                const float err = inv - tg[0];
                // confidence target: a valid target the proposal lies within
                // conf_tol of (relative, in inverse depth)
                const bool good = v && fabsf(err) <= a.conf_tol * fmaxf(tg[0], 1e-3f);
                if (v) {
                    ld = fabsf(err);
                    go[0] = a.w_depth * (err > 0 ? 1.f : (err < 0 ? -1.f : 0.f)) * inv * (1.f - inv);
                    const float dot = nx * tg[1] + ny * tg[2] + nz * tg[3];
                    ln = 1.0f - dot;
                    // d(1 - n.t)/do = -(t - (n.t) n) / |o|
                    go[1] = -a.w_normal * (tg[1] - dot * nx) / len;
                    go[2] = -a.w_normal * (tg[2] - dot * ny) / len;
                    go[3] = -a.w_normal * (tg[3] - dot * nz) / len;
                }
                const float ct = good ? 1.f : 0.f;
                lc = -(ct * __logf(fmaxf(conf, 1e-7f)) + (1.f - ct) * __logf(fmaxf(1.f - conf, 1e-7f)));
                go[4] = a.w_conf * (conf - ct) * a.scale_c;
                for (int k = 0; k < 4; ++k) go[k] *= a.scale_v;
            }
            float* gf = a.gfeat + (size_t)p * 32;
            for (int c = 0; c < 32; ++c) {
                float s = 0.f;
                for (int k = 0; k < 5; ++k) s += s_w[k * 32 + c] * go[k];
                gf[c] = s;
            }
            // per-wave partial sums of dW, db and the loss terms
            const int wv = threadIdx.x >> 6;
            for (int k = 0; k < 5; ++k) {
                for (int c = 0; c < 32; ++c) atomicAdd(&s_red[wv][k * 32 + c], go[k] * f[c]);
                atomicAdd(&s_red[wv][5 * 32 + k], go[k]);
            }
            if (inside) {
                atomicAdd(&s_red[wv][5 * 32 + 5], ld);
                atomicAdd(&s_red[wv][5 * 32 + 6], ln);
                atomicAdd(&s_red[wv][5 * 32 + 7], lc);
                const float* tg = a.target + 5 * (((size_t)n * a.h_valid + y) * a.w_valid + x);
                atomicAdd(&s_red[wv][5 * 32 + 8], tg[4] > 0.5f ? 1.f : 0.f);
            }
        }
    }
    __syncthreads();
    if (a.gfeat) {
        for (int i = threadIdx.x; i < 5 * 32 + 9; i += blockDim.x) {
            float s = 0.f;
            for (int wv = 0; wv < (int)(blockDim.x >> 6); ++wv) s += s_red[wv][i];
            if (i < 5 * 32 + 5) atomicAdd(&a.gw[i], s);
            else atomicAdd(&a.loss[i - (5 * 32 + 5)], s);
        }
    }
}

// --------------------------------------------------------------------------
// Adam with fp32 master weights; refreshes the bf16 copies
// --------------------------------------------------------------------------

__global__ void adam_kernel(AdamArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const float g = a.g[i] * a.gscale;
    const float m = a.b1 * a.m[i] + (1.f - a.b1) * g;
    const float v = a.b2 * a.v[i] + (1.f - a.b2) * g * g;
    a.m[i] = m;
    a.v[i] = v;
    const float mh = m / a.c1, vh = v / a.c2;
    a.p[i] -= a.lr * (mh / (sqrtf(vh) + a.eps) + a.wd * a.p[i]);
    a.g[i] = 0.0f;
}

// master [tap][ci][co] fp32 -> forward copy [tap][co][ci] and flipped data-gradient
// copy [8 - tap][ci][co] (bf16); ci beyond cin_used are zero in both
__global__ void wprep_kernel(const float* __restrict__ w, int cin, int cout, uint16_t* __restrict__ wf,
                             uint16_t* __restrict__ wb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)9 * cin * cout;
    if (i >= total) return;
    const int co = (int)(i % cout);
    const int ci = (int)((i / cout) % cin);
    const int t = (int)(i / ((int64_t)cin * cout));
    const uint16_t b = f2bf(w[i]);
    wf[((size_t)t * cout + co) * cin + ci] = b;
    if (wb) wb[((size_t)(8 - t) * cin + ci) * cout + co] = b;
}

// --------------------------------------------------------------------------
// launchers
// --------------------------------------------------------------------------

static inline unsigned nb(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_conv(const ConvArgs& a, int N, hipStream_t st) {
    if (a.mode == 0) {
        const int tiles = ((a.W + CT - 1) / CT) * ((a.H + CT - 1) / CT);
        hipLaunchKernelGGL(conv3x3_lds_kernel, dim3(tiles, a.cout / 32, N), dim3(256), 0, st, a);
    } else {
        const int tiles = ((a.W + 15) / 16) * ((a.H + 7) / 8);
        hipLaunchKernelGGL(conv3x3_kernel, dim3(tiles, a.cout / 32, N), dim3(256), 0, st, a);
    }
    return hipGetLastError();
}

hipError_t launch_wgrad(const WgradArgs& a, hipStream_t st) {
    const int64_t npix = (int64_t)a.N * a.H * a.W;
    const unsigned chunks = (unsigned)((npix + a.chunk - 1) / a.chunk);
    hipLaunchKernelGGL(wgrad_kernel, dim3(chunks, ((a.c0 + a.c1) / 32) * (a.cout / 32), 9), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_bgrad(const uint16_t* dy, int64_t npix, int cout, float* db, hipStream_t st) {
    const int64_t chunk = 8192;
    hipLaunchKernelGGL(bgrad_kernel, dim3((unsigned)((npix + chunk - 1) / chunk)), dim3(256), 0, st, dy, npix, cout,
                       chunk, db);
    return hipGetLastError();
}

hipError_t launch_relu_bwd(const float* g, const uint16_t* y, int64_t n, uint16_t* out, hipStream_t st) {
    hipLaunchKernelGGL(relu_bwd_kernel, dim3(nb(n)), dim3(256), 0, st, g, y, n, out);
    return hipGetLastError();
}

hipError_t launch_pool_fwd(const uint16_t* x, int N, int H, int W, int C, uint16_t* y, hipStream_t st) {
    hipLaunchKernelGGL(pool_fwd_kernel, dim3(nb((int64_t)N * (H / 2) * (W / 2) * C)), dim3(256), 0, st, x, N, H, W, C,
                       y);
    return hipGetLastError();
}

hipError_t launch_pool_bwd(const float* gy, int N, int H, int W, int C, float* gx, hipStream_t st) {
    hipLaunchKernelGGL(pool_bwd_kernel, dim3(nb((int64_t)N * H * W * C)), dim3(256), 0, st, gy, N, H, W, C, gx);
    return hipGetLastError();
}

hipError_t launch_zero(float* p, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_f32_kernel, dim3(nb(n)), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

hipError_t launch_head(const HeadArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(head_kernel, dim3(nb(a.npix)), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_adam(const AdamArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(adam_kernel, dim3(nb(a.n)), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_wprep(const float* w, int cin, int cout, uint16_t* wf, uint16_t* wb, hipStream_t st) {
    hipLaunchKernelGGL(wprep_kernel, dim3(nb((int64_t)9 * cin * cout)), dim3(256), 0, st, w, cin, cout, wf, wb);
    return hipGetLastError();
}

}  // namespace net
}  // namespace mvs
